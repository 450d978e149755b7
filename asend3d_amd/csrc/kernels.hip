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
// brick8 element integration + assembly (FP64).
//   K_ab,ij = sum_gp w detJ [lam g_a,i g_b,j + mu g_a,j g_b,i + mu (g_a.g_b) d_ij]
// is B_a^T C B_b for isotropic C (Lame lam, mu), the same operator as
// get_ruk_dfd0's B^T C B (S/element_equations.rs:398-411), with g_a the
// physical gradient of shape function a at Gauss point gp (8 points,
// +-1/sqrt(3), weight 1).
// ---------------------------------------------------------------------------
__constant__ double kHexNdSign[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                        {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

// natural derivatives dN_t/ds_c of the brick8 shape functions at the 8 Gauss
// points (+-1/sqrt(3), point gp has s_c = +g when bit c of gp is set)
struct HexDN {
    double d[8][8][3];
};
constexpr HexDN make_hex_dn() {
    HexDN h{};
    const double g = 0.5773502691896257;
    const double sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                             {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    for (int gp = 0; gp < 8; ++gp) {
        const double p0 = (gp & 1) ? g : -g, p1 = (gp & 2) ? g : -g, p2 = (gp & 4) ? g : -g;
        for (int t = 0; t < 8; ++t) {
            const double fx = 1.0 + sg[t][0] * p0, fy = 1.0 + sg[t][1] * p1, fz = 1.0 + sg[t][2] * p2;
            h.d[gp][t][0] = 0.125 * sg[t][0] * fy * fz;
            h.d[gp][t][1] = 0.125 * fx * sg[t][1] * fz;
            h.d[gp][t][2] = 0.125 * fx * fy * sg[t][2];
        }
    }
    return h;
}
__constant__ HexDN kHexDN = make_hex_dn();

// node sign pattern of the brick8 reference element (kHexNdSign by index bits)
__device__ __forceinline__ void hex_sign(int t, double& sx, double& sy, double& sz) {
    sx = (((t ^ (t >> 1)) & 1) != 0) ? 1.0 : -1.0;
    sy = (t & 2) ? 1.0 : -1.0;
    sz = (t & 4) ? 1.0 : -1.0;
}

// two doubles as a clang vector (the form the non-temporal builtins take)
typedef double d2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double group8_sum(double v) {
    v += __shfl_xor(v, 1, kWave);
    v += __shfl_xor(v, 2, kWave);
    v += __shfl_xor(v, 4, kWave);
    return v;
}

// Element-once brick8 integration for general (non-translated) meshes: one
// wavefront per element, 4 per workgroup.  Phase 1, lane = 8 gp + t: the
// Jacobian at Gauss point gp by an 8-lane butterfly over the nodes t, then
// the physical gradient g_t, N_t and detJ staged in LDS.  Phase 2, lane =
// 8 a + b with a <= b: G_ab = sum_gp detJ g_a g_b^T, and for isotropic C the
// block is K_ab = lam G_ab + mu G_ab^T + mu tr(G_ab) I, plus c_mass rho M_ab I
// and, with c_damp != 0, c_damp times the same integral for a general 6x6
// material damping matrix; with Cs (a general 6x6 material stiffness per
// element, orthotropic materials) k_scale E_e times that integral replaces
// the isotropic part; K_ba = K_ab^T.  With kcond (6 conductivities per
// element, k11 k22 k33 k12 k13 k23) the block is the heat conduction term
// s_ab I with s_ab = sum over m, n of k_mn G_ab[m][n] = int g_a^T k g_b: the
// scalar conduction matrix on each of the three DOFs of a node, of which the
// first carries the temperature.  The element's eight row records are staged
// in LDS and written as coalesced 16-byte stores.
__global__ __launch_bounds__(256) void k_hex8_elem_rec(int n_el, const int32_t* __restrict__ conn_r,
                                                       const double* __restrict__ xyz_r,
                                                       const double* __restrict__ E, const double* __restrict__ nu,
                                                       const double* __restrict__ rho, double k_scale, double c_mass,
                                                       const double* __restrict__ Cd,
                                                       const int32_t* __restrict__ cd_idx, double c_damp,
                                                       const double* __restrict__ kcond,
                                                       const double* __restrict__ Cs,
                                                       const int32_t* __restrict__ cs_idx,
                                                       double* __restrict__ Kr) {
    __shared__ __align__(16) double s_g[4][8][8][4];  // (g_t, N_t) at each point
    __shared__ double s_det[4][8];
    __shared__ __align__(16) double s_out[4][8 * kRecStride];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int e = blockIdx.x * 4 + wv;
    const bool live = e < n_el;
    const int ec = live ? e : n_el - 1;
    {
        const int gp = lane >> 3, t = lane & 7;
        const double g = 0.5773502691896257;
        const int node = conn_r[8 * (int64_t)ec + t];
        double X[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) X[c] = xyz_r[3 * (int64_t)node + c];
        double sx, sy, sz;
        hex_sign(t, sx, sy, sz);
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
        double4 gn;
        gn.x = dn[0] * Ji[0] + dn[1] * Ji[3] + dn[2] * Ji[6];
        gn.y = dn[0] * Ji[1] + dn[1] * Ji[4] + dn[2] * Ji[7];
        gn.z = dn[0] * Ji[2] + dn[1] * Ji[5] + dn[2] * Ji[8];
        gn.w = 0.125 * fx * fy * fz;
        *reinterpret_cast<double4*>(s_g[wv][gp][t]) = gn;
        if (t == 0) s_det[wv][gp] = det;
    }
    __syncthreads();

    const int a = lane >> 3, b = lane & 7;
    if (a <= b) {
        double G[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) G[c] = 0.0;
        double m = 0.0;
#pragma unroll 2
        for (int q = 0; q < 8; ++q) {
            const double dq = s_det[wv][q];
            const double4 ga = *reinterpret_cast<const double4*>(s_g[wv][q][a]);
            const double4 gb = *reinterpret_cast<const double4*>(s_g[wv][q][b]);
            const double da[3] = {dq * ga.x, dq * ga.y, dq * ga.z};
            const double gbv[3] = {gb.x, gb.y, gb.z};
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) G[3 * i + j] += da[i] * gbv[j];
            m += dq * ga.w * gb.w;
        }
        const double Ee = k_scale * E[ec], ne = nu[ec];
        // a general material matrix (Cs) replaces the isotropic one, weighted by k_scale E_e
        const double lam = Cs ? 0.0 : Ee * ne / ((1.0 + ne) * (1.0 - 2.0 * ne));
        const double mu = Cs ? 0.0 : Ee / (2.0 * (1.0 + ne));
        const double dg = mu * (G[0] + G[4] + G[8]) + (c_mass != 0.0 ? c_mass * rho[ec] * m : 0.0);
        // c_damp B^T C_d B for a general symmetric 6x6 matrix in the strain
        // order (11, 22, 33, 12, 13, 23), engineering shear: block entry
        // (i, j) = sum over m, n of C_d[V(i,m)][V(j,n)] G_mn
        double X[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) X[c] = 0.0;
        // w sum over m, n of Dm[V(i,m)][V(j,n)] G_mn into X (strain order 11, 22, 33, 12, 13, 23)
        auto general_block = [&](const double* Dm, double w) {
            constexpr int V[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    double acc = 0.0;
#pragma unroll
                    for (int m2 = 0; m2 < 3; ++m2)
#pragma unroll
                        for (int n2 = 0; n2 < 3; ++n2) acc += Dm[6 * V[i][m2] + V[j][n2]] * G[3 * m2 + n2];
                    X[3 * i + j] += w * acc;
                }
        };
        if (c_damp != 0.0) general_block(Cd + 36 * (cd_idx ? (int64_t)cd_idx[ec] : 0), c_damp);
        if (Cs && Ee != 0.0) general_block(Cs + 36 * (cs_idx ? (int64_t)cs_idx[ec] : 0), Ee);
        double* o = &s_out[wv][kRecStride * a + 9 * b];
        double* ot = &s_out[wv][kRecStride * b + 9 * a];
        if (kcond) {
            const double* kc = kcond + 6 * (int64_t)ec;
            // conduction (k_scale) plus heat capacity c_mass rho_e int N_a N_b, rho_e = density x specific heat
            const double s = k_scale * (kc[0] * G[0] + kc[1] * G[4] + kc[2] * G[8] + kc[3] * (G[1] + G[3]) +
                                        kc[4] * (G[2] + G[6]) + kc[5] * (G[5] + G[7])) +
                             (c_mass != 0.0 ? c_mass * rho[ec] * m : 0.0);
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                o[c] = (c % 4 == 0) ? s : 0.0;
                if (a != b) ot[c] = (c % 4 == 0) ? s : 0.0;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const double v = lam * G[3 * i + j] + mu * G[3 * j + i] + (i == j ? dg : 0.0) + X[3 * i + j];
                    o[3 * i + j] = v;
                    if (a != b) ot[3 * j + i] = v;
                }
        }
    }
    __syncthreads();
    if (!live) return;
    // the records are streamed out (1.47 GB on a 320,000-element mesh, far
    // beyond the caches): non-temporal stores, 5% faster than plain ones
    const d2v* src = reinterpret_cast<const d2v*>(s_out[wv]);
    d2v* dst = reinterpret_cast<d2v*>(Kr + (int64_t)8 * kRecStride * e);
    for (int h = lane; h < 4 * kRecStride; h += kWave) __builtin_nontemporal_store(src[h], dst + h);
}

// Row gather of element records into the sliced-ELL operator.  A 128-thread
// workgroup owns 8 rows of a 64-row slice (16 lanes per row).  Per chunk of
// up to 8 records of each row (n2rec lists record 8 e + a of every element e
// at the row's node, a = the node's local index), phase 1 copies the records
// into LDS, two lanes per record with sequential 16-byte loads (every fetched
// line is used whole, once), together with the records' in-row slot indices
// (slot[8 rec + b] = slot of block b).  Phase 2: lane j of a row handles slots
// j and j + 16 of each 32-slot pass and adds, in record order, the block of
// every staged record that maps to the slot.  Phase 3 transposes the sums
// through LDS so that each (slot, component) of the 8 rows is one 64-byte
// store; every value (padding included) is written once.  NL adds the
// internal force of the row node from the element residuals
// Re[24 e + 3 a + i] and, when rhs is set, rhs = fact f - rint.  Workgroups
// are visited in XCD-contiguous ranges.
constexpr int kRgRows = 8;   // rows per workgroup
constexpr int kRgLanes = 16; // lanes per row
constexpr int kRgThreads = kRgRows * kRgLanes;

template <bool NL>
__global__ __launch_bounds__(kRgThreads) void k_rec_gather(
    int n_rows, int n_slices, const int32_t* __restrict__ sptr, const int32_t* __restrict__ n2e_ptr,
    const int32_t* __restrict__ n2rec, const uint8_t* __restrict__ slot, const double* __restrict__ Kr,
    const double* __restrict__ Re, double fact, const double* __restrict__ f, double* __restrict__ K,
    double* __restrict__ rint, double* __restrict__ rhs) {
    // records of the chunk, then (phase 3) the pass's sums [slot][9][row]
    __shared__ __align__(16) double s_rec[kRgRows * 8 * kRecStride];
    __shared__ uint64_t s_sm[kRgRows][8];
    static_assert(2 * kRgLanes * 9 * kRgRows <= kRgRows * 8 * kRecStride, "phase-3 buffer");
    static_assert(kRecStride % 4 == 0, "two halves of 16-byte pieces");
    constexpr int kHalf = kRecStride / 2;  // doubles per half record
    const int g = xcd_block(blockIdx.x, gridDim.x);
    const int s = g / (64 / kRgRows), part = g % (64 / kRgRows);
    if (s >= n_slices) return;
    const int rl = threadIdx.x / kRgLanes, j = threadIdx.x % kRgLanes;
    const int lr0 = part * kRgRows;  // first row of the workgroup in the slice
    const int r = 64 * s + lr0 + rl;
    const bool live = r < n_rows;
    const int s0 = sptr[s], w = sptr[s + 1] - s0;
    const int q0 = live ? n2e_ptr[r] : 0, nq = live ? n2e_ptr[r + 1] - q0 : 0;
    if (NL) {
        double ri[3] = {0.0, 0.0, 0.0};
        for (int q = j; q < nq; q += kRgLanes) {
            const int rec = n2rec[q0 + q];
#pragma unroll
            for (int i = 0; i < 3; ++i) ri[i] += Re[24 * (int64_t)(rec >> 3) + 3 * (rec & 7) + i];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int o = 1; o < kRgLanes; o <<= 1) ri[i] += __shfl_xor(ri[i], o, kWave);
        if (live && j == 0) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const int64_t d = 3 * (int64_t)r + i;
                rint[d] = ri[i];
                if (rhs) rhs[d] = fact * f[d] - ri[i];
            }
        }
    }
    const double* my_rec = s_rec + rl * 8 * kRecStride;
    constexpr uint64_t kOnes = 0x0101010101010101ull, kHighs = 0x8080808080808080ull;
    for (int kb = 0; kb < w; kb += 2 * kRgLanes) {
        double acc[2][9];
#pragma unroll
        for (int sk = 0; sk < 2; ++sk)
#pragma unroll
            for (int c = 0; c < 9; ++c) acc[sk][c] = 0.0;
        for (int c0 = 0;; c0 += 8) {
            if (c0 > 0 || kb > 0) __syncthreads();
            // phase 1: lanes 2 qi + h copy half h of record c0 + qi of the row
            // with sequential 16-byte loads (every fetched line is used whole)
            const int qi = j >> 1, h = j & 1;
            const int q = c0 + qi;
            if (q < nq) {
                const int rec = n2rec[q0 + q];
                if (h == 0) s_sm[rl][qi] = *reinterpret_cast<const uint64_t*>(slot + 8 * (int64_t)rec);
                const double2* src =
                    reinterpret_cast<const double2*>(Kr + (int64_t)kRecStride * rec + kHalf * h);
                double2* dst = reinterpret_cast<double2*>(s_rec + (rl * 8 + qi) * kRecStride + kHalf * h);
                // plain loads: non-temporal ones measured 1.9x slower here
#pragma unroll
                for (int v = 0; v < kHalf / 2; ++v) dst[v] = src[v];
            } else if (h == 0) {
                s_sm[rl][qi] = ~0ull;  // slot 255 is never used
            }
            const int more = __syncthreads_or(c0 + 8 < nq);
            // phase 2
#pragma unroll
            for (int sk = 0; sk < 2; ++sk) {
                const int k = kb + j + kRgLanes * sk;
                if (k >= w) continue;
                const uint64_t kk = kOnes * (uint64_t)k;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const uint64_t x = s_sm[rl][u] ^ kk;
                    const uint64_t z = (x - kOnes) & ~x & kHighs;  // lowest set bit marks the matching byte
                    if (z) {
                        const double* blk = my_rec + u * kRecStride + 9 * (__builtin_ctzll(z) >> 3);
#pragma unroll
                        for (int c = 0; c < 9; ++c) acc[sk][c] += blk[c];
                    }
                }
            }
            if (!more) break;
        }
        // phase 3: sums to LDS as [slot in pass][component][row], then runs of
        // 8 rows to K
        __syncthreads();
#pragma unroll
        for (int sk = 0; sk < 2; ++sk)
#pragma unroll
            for (int c = 0; c < 9; ++c) s_rec[((j + kRgLanes * sk) * 9 + c) * kRgRows + rl] = acc[sk][c];
        __syncthreads();
        const int np = min(2 * kRgLanes, w - kb);
        for (int idx = threadIdx.x; idx < np * 9 * kRgRows; idx += kRgThreads) {
            const int row = idx % kRgRows, kc = idx / kRgRows;
            K[(int64_t)9 * 64 * (s0 + kb + kc / 9) + 64 * (kc % 9) + lr0 + row] = s_rec[idx];
        }
    }
}

// Reference blocks of one brick8 element with E = 1 and rho = 1: lane (a, b)
// writes Khat[9 (8a + b) + 3i + j] = K_ab,ij / E and Mhat[8a + b] =
// sum_gp detJ N_a N_b, by the same arithmetic as the general kernel.  On a
// mesh whose elements are all translates of this one with one Poisson ratio,
// K_e = E_e Khat and M_e = rho_e Mhat exactly (up to rounding).
__global__ __launch_bounds__(64) void k_hex8_reference(const int32_t* __restrict__ conn_r,
                                                       const double* __restrict__ xyz_r, double nu,
                                                       double* __restrict__ Khat, double* __restrict__ Mhat) {
    const int lane = threadIdx.x;
    const int a = lane / 8, b = lane % 8;
    double X[8][3];
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int c = 0; c < 3; ++c) X[t][c] = xyz_r[3 * (int64_t)conn_r[t] + c];
    const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = 1.0 / (2.0 * (1.0 + nu));
    double sa[3], sb[3];
    hex_sign(a, sa[0], sa[1], sa[2]);
    hex_sign(b, sb[0], sb[1], sb[2]);
    const double g = 0.5773502691896257;
    double kb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double m = 0.0;
    for (int gp = 0; gp < 8; ++gp) {
        const double p0 = (gp & 1) ? g : -g, p1 = (gp & 2) ? g : -g, p2 = (gp & 4) ? g : -g;
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int t = 0; t < 8; ++t) {
            const double d0 = kHexDN.d[gp][t][0], d1 = kHexDN.d[gp][t][1], d2 = kHexDN.d[gp][t][2];
            for (int i = 0; i < 3; ++i) {
                J[i][0] += X[t][i] * d0;
                J[i][1] += X[t][i] * d1;
                J[i][2] += X[t][i] * d2;
            }
        }
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        const double id = 1.0 / det;
        const double Ji[9] = {c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id,
                              (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id, c01 * id,
                              (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id, (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id,
                              c02 * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id,
                              (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id};
        const double fxa = 1.0 + sa[0] * p0, fya = 1.0 + sa[1] * p1, fza = 1.0 + sa[2] * p2;
        const double fxb = 1.0 + sb[0] * p0, fyb = 1.0 + sb[1] * p1, fzb = 1.0 + sb[2] * p2;
        const double da[3] = {0.125 * sa[0] * fya * fza, 0.125 * fxa * sa[1] * fza, 0.125 * fxa * fya * sa[2]};
        const double db[3] = {0.125 * sb[0] * fyb * fzb, 0.125 * fxb * sb[1] * fzb, 0.125 * fxb * fyb * sb[2]};
        double ga[3], gb[3];
        for (int c = 0; c < 3; ++c) {
            ga[c] = da[0] * Ji[0 + c] + da[1] * Ji[3 + c] + da[2] * Ji[6 + c];
            gb[c] = db[0] * Ji[0 + c] + db[1] * Ji[3 + c] + db[2] * Ji[6 + c];
        }
        const double dot = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                kb[3 * i + j] += det * (lam * ga[i] * gb[j] + mu * ga[j] * gb[i] + (i == j ? mu * dot : 0.0));
        m += det * (0.125 * fxa * fya * fza) * (0.125 * fxb * fyb * fzb);
    }
    for (int c = 0; c < 9; ++c) Khat[9 * lane + c] = kb[c];
    Mhat[lane] = m;
}

// Row-wise brick8 / brickIM assembly for meshes of translated copies of one
// element: K_row,slot = sum over the elements e holding both the row node and
// the slot node of k_scale E_e Khat[a][b] (+ c_mass rho_e Mhat[a][b] I).  The (at most 8)
// elements of the row node and their connectivity are held in registers; the
// stores are the same coalesced slot-by-slot writes as the general kernel.
__global__ __launch_bounds__(256) void k_hex8_assemble_uniform(
    int n_rows, int n_slices, const int32_t* __restrict__ sptr, const int32_t* __restrict__ scol,
    const int32_t* __restrict__ rp, const int32_t* __restrict__ n2e_ptr, const int32_t* __restrict__ n2e,
    const int32_t* __restrict__ conn_r, const double* __restrict__ Khat, const double* __restrict__ Mhat,
    const double* __restrict__ E, const double* __restrict__ rho, double k_scale, double c_mass,
    double* __restrict__ K) {
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (s >= n_slices) return;
    const int r = 64 * s + lane;
    const bool live = r < n_rows;
    const int s0 = sptr[s], w = sptr[s + 1] - s0;
    const int nrow = live ? rp[r + 1] - rp[r] : 0;
    const int q0 = live ? n2e_ptr[r] : 0, nq = live ? n2e_ptr[r + 1] - q0 : 0;
    int cn[8][8], av[8];
    double ev[8], mv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        av[q] = 0;
        ev[q] = mv[q] = 0.0;
#pragma unroll
        for (int t = 0; t < 8; ++t) cn[q][t] = -1;
        if (q < nq) {
            const int e = n2e[q0 + q];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                cn[q][t] = conn_r[8 * (int64_t)e + t];
                if (cn[q][t] == r) av[q] = t;
            }
            ev[q] = k_scale * E[e];
            mv[q] = c_mass * rho[e];
        }
    }
    for (int k = 0; k < w; ++k) {
        double acc[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) acc[c] = 0.0;
        if (k < nrow) {
            const int col = scol[64 * (int64_t)(s0 + k) + lane];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                int bq = -1;
#pragma unroll
                for (int t = 0; t < 8; ++t) bq = (cn[q][t] == col) ? t : bq;
                if (bq >= 0) {
                    const int ab = 8 * av[q] + bq;
                    for (int c = 0; c < 9; ++c) acc[c] += ev[q] * Khat[9 * ab + c];
                    const double mm = mv[q] * Mhat[ab];
                    acc[0] += mm;
                    acc[4] += mm;
                    acc[8] += mm;
                }
            }
        }
        double* base = K + (int64_t)9 * 64 * (s0 + k) + lane;
#pragma unroll
        for (int c = 0; c < 9; ++c) base[64 * c] = acc[c];
    }
}

void launch_hex8_elem_rec(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                          const double* nu, const double* rho, double k_scale, double c_mass, const double* Cd,
                          const int32_t* cd_idx, double c_damp, const double* kcond, const double* Cs,
                          const int32_t* cs_idx, double* Kr) {
    if (n_el <= 0) return;
    hipLaunchKernelGGL(k_hex8_elem_rec, dim3((unsigned)((n_el + 3) / 4)), dim3(256), 0, st, n_el, conn_r, xyz_r, E,
                       nu, rho, k_scale, c_mass, Cd, cd_idx, c_damp, kcond, Cs, cs_idx, Kr);
}

void launch_rec_gather(hipStream_t st, int n_rows, int n_slices, const int32_t* sptr, const int32_t* n2e_ptr,
                       const int32_t* n2rec, const uint8_t* slot, const double* Kr, const double* Re, double fact,
                       const double* f, double* K, double* rint, double* rhs) {
    if (n_slices <= 0) return;
    const dim3 grid((unsigned)(n_slices * (64 / kRgRows))), block(kRgThreads);
    if (Re)
        hipLaunchKernelGGL(k_rec_gather<true>, grid, block, 0, st, n_rows, n_slices, sptr, n2e_ptr, n2rec, slot, Kr,
                           Re, fact, f, K, rint, rhs);
    else
        hipLaunchKernelGGL(k_rec_gather<false>, grid, block, 0, st, n_rows, n_slices, sptr, n2e_ptr, n2rec, slot, Kr,
                           nullptr, 0.0, nullptr, K, nullptr, nullptr);
}

void launch_hex8_reference(hipStream_t st, const int32_t* conn_r, const double* xyz_r, double nu, double* Khat,
                           double* Mhat) {
    hipLaunchKernelGGL(k_hex8_reference, dim3(1), dim3(64), 0, st, conn_r, xyz_r, nu, Khat, Mhat);
}

void launch_hex8_assemble_uniform(hipStream_t st, int n_rows, int n_slices, const int32_t* sptr, const int32_t* scol,
                                  const int32_t* rp, const int32_t* n2e_ptr, const int32_t* n2e,
                                  const int32_t* conn_r, const double* Khat, const double* Mhat, const double* E,
                                  const double* rho, double k_scale, double c_mass, double* K) {
    if (n_slices <= 0) return;
    hipLaunchKernelGGL(k_hex8_assemble_uniform, dim3((unsigned)((n_slices + 3) / 4)), dim3(256), 0, st, n_rows,
                       n_slices, sptr, scol, rp, n2e_ptr, n2e, conn_r, Khat, Mhat, E, rho, k_scale, c_mass, K);
}

// ---------------------------------------------------------------------------
// Grid reductions: every block writes its partial sums with plain stores;
// a single-block finalize kernel (stream-ordered after the producer, so the
// partials are visible without device-scope fences) adds them in a fixed
// order and updates the loop scalars.  Deterministic run to run.
// ---------------------------------------------------------------------------
enum FinOp {
    FIN_PQ = 0,
    FIN_RR = 1,
    FIN_INIT = 2,
    FIN_UPDATE = 3,
    FIN_DOT = 4,
    FIN_MAX = 5,
    FIN_CG1_INIT = 7,
    FIN_CG1 = 8,
    FIN_RU_RES = 9,
    FIN_CG1_RU = 10,
    FIN_PQR = 11,
    FIN_FUPD = 12
};

__global__ __launch_bounds__(kBlock) void k_finalize(const double* __restrict__ partials, int nparts, int op,
                                                     Scalars* sc) {
    if ((op == FIN_PQ || op == FIN_UPDATE || op == FIN_CG1) && sc->done) return;
    if (op == FIN_RU_RES && !sc->gate) return;
    if (op == FIN_CG1_RU && (sc->done || sc->gate)) return;
    if ((op == FIN_PQR || op == FIN_FUPD) && sc->done) return;
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
    const int nv = (op == FIN_INIT || op == FIN_UPDATE || op == FIN_PQR || op == FIN_FUPD)
                       ? 2
                       : ((op == FIN_CG1 || op == FIN_CG1_RU) ? 3 : 1);
    double t[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        t[0] += partials[(size_t)nv * b];
        if (nv >= 2) t[1] += partials[(size_t)nv * b + 1];
        if (nv == 3) t[2] += partials[(size_t)nv * b + 2];
    }
    block_sum<3>(t);
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
        case FIN_CG1_INIT:
            // single-reduction PCG start: t0 = |b|^2
            sc->ff = t[0];
            sc->rz = 0.0;
            sc->it = 0;
            sc->done = t[0] == 0.0;
            break;
        case FIN_CG1: {
            // single-reduction PCG: t0 = gamma = r.u, t1 = delta = (K u).u,
            // t2 = r.r of the current residual
            const double g = t[0], d = t[1];
            sc->rr = t[2];
            const bool conv = sc->stop_abs ? (g <= sc->tol) : (t[2] <= sc->tol2 * sc->ff);
            if (conv || sc->it >= sc->max_it) {
                sc->done = 1;
                break;
            }
            const double beta = sc->it == 0 ? 0.0 : g / sc->rz;
            const double den = sc->it == 0 ? d : d - beta * g / sc->alpha;
            if (!(den > 0.0)) {  // lost positive definiteness: stop, reported unconverged
                sc->done = 1;
                break;
            }
            sc->beta = beta;
            sc->alpha = g / den;
            sc->rz = g;
            sc->it += 1;
            break;
        }
        case FIN_RU_RES:
            // reliable update: the FP64 residual decides convergence
            sc->rr = t[0];
            sc->n_resid += 1;
            if (t[0] <= sc->tol2 * sc->ff || sc->n_resid >= kMixedMaxRefresh || sc->it >= sc->max_it) {
                sc->done = 1;
                sc->gate = 0;
            }
            break;
        case FIN_CG1_RU: {
            // FP32 single-reduction step; sc->rr is the FP64 residual of the
            // last refresh, sc->zq the refresh factor squared
            const double g = t[0], d = t[1];
            sc->pq = t[2];  // recursive residual, FP32 recurrence
            if (t[2] <= sc->zq * sc->rr || t[2] <= sc->tol2 * sc->ff) {
                sc->gate = 1;
                break;
            }
            if (sc->it >= sc->max_it) {
                sc->done = 1;
                break;
            }
            const double beta = sc->it == 0 ? 0.0 : g / sc->rz;
            const double den = sc->it == 0 ? d : d - beta * g / sc->alpha;
            if (!(den > 0.0)) {
                sc->done = 1;
                break;
            }
            sc->beta = beta;
            sc->alpha = g / den;
            sc->rz = g;
            sc->it += 1;
            break;
        }
        case FIN_PQR:
            // flexible step length: t0 = p.q, t1 = p.r
            sc->pq = t[0];
            sc->alpha = t[1] / t[0];
            break;
        case FIN_FUPD: {
            // flexible direction: t0 = (Dinv r).q, t1 = r.r
            sc->beta = -t[0] / sc->pq;
            sc->rr = t[1];
            sc->it += 1;
            if (t[1] <= sc->tol2 * sc->ff || sc->it >= sc->max_it || !(t[1] == t[1])) sc->done = 1;
            break;
        }
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
template <bool NT, typename T>
__device__ __forceinline__ T sell_load(const T* p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <typename T, int MODE, bool NT>
__global__ __launch_bounds__(kSellMaxWaves * kWave) void k_sell3(int n_rows, int n_slices,
                                                                 const int32_t* __restrict__ sptr,
                                                  const int32_t* __restrict__ scol, const T* __restrict__ val,
                                                  const T* __restrict__ x, T* __restrict__ y,
                                                  const T* __restrict__ f, double* partials, const Scalars* sc) {
    if ((MODE == 1 || MODE == 4 || MODE == 7) && sc->done) return;
    if (MODE == 6 && (sc->done || sc->gate)) return;
    if (MODE == 5 && !sc->gate) return;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int slice = lb * (int)(blockDim.x / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int r = slice * kWave + lane;
    double part = 0.0, part1 = 0.0, part2 = 0.0;
    if (slice < n_slices) {
        const int k0 = sptr[slice], k1 = sptr[slice + 1];
        const int32_t* cp = scol + (int64_t)k0 * kWave + lane;
        const T* vp = val + (int64_t)k0 * 9 * kWave + lane;
        T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll 2
        for (int k = k0; k < k1; ++k) {
            // values and columns of operators larger than the Infinity Cache
            // are streamed once: non-temporal loads keep them from evicting
            // the gathered x from L2
            const int c = sell_load<NT>(cp);
            const T x0 = x[3 * (int64_t)c], x1 = x[3 * (int64_t)c + 1], x2 = x[3 * (int64_t)c + 2];
            T v[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) v[q] = sell_load<NT>(vp + q * kWave);
            a0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
            a1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
            a2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
            cp += kWave;
            vp += 9 * kWave;
        }
        if (r < n_rows) {
            if (MODE == 2 || MODE == 3 || MODE == 5) {
                T y0 = f[3 * (int64_t)r] - a0, y1 = f[3 * (int64_t)r + 1] - a1, y2 = f[3 * (int64_t)r + 2] - a2;
                y[3 * (int64_t)r] = y0;
                y[3 * (int64_t)r + 1] = y1;
                y[3 * (int64_t)r + 2] = y2;
                part = (double)y0 * y0 + (double)y1 * y1 + (double)y2 * y2;
            } else if (MODE == 7) {
                // x = p, f = r: partials p.(K p) and p.r
                y[3 * (int64_t)r] = a0;
                y[3 * (int64_t)r + 1] = a1;
                y[3 * (int64_t)r + 2] = a2;
                const double p0 = x[3 * (int64_t)r], p1 = x[3 * (int64_t)r + 1], p2 = x[3 * (int64_t)r + 2];
                part = (double)a0 * p0 + (double)a1 * p1 + (double)a2 * p2;
                part1 = p0 * f[3 * (int64_t)r] + p1 * f[3 * (int64_t)r + 1] + p2 * f[3 * (int64_t)r + 2];
            } else if (MODE == 4 || MODE == 6) {
                // x = u (preconditioned residual), f = r: partials r.u, (K u).u, r.r
                y[3 * (int64_t)r] = a0;
                y[3 * (int64_t)r + 1] = a1;
                y[3 * (int64_t)r + 2] = a2;
                const double u0 = x[3 * (int64_t)r], u1 = x[3 * (int64_t)r + 1], u2 = x[3 * (int64_t)r + 2];
                const double r0 = f[3 * (int64_t)r], r1 = f[3 * (int64_t)r + 1], r2 = f[3 * (int64_t)r + 2];
                part = r0 * u0 + r1 * u1 + r2 * u2;
                part1 = (double)a0 * u0 + (double)a1 * u1 + (double)a2 * u2;
                part2 = r0 * r0 + r1 * r1 + r2 * r2;
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
    if (MODE == 4 || MODE == 6) {
        double t[3] = {part, part1, part2};
        block_partials<3>(t, partials);
        return;
    }
    if (MODE == 7) {
        double t[2] = {part, part1};
        block_partials<2>(t, partials);
        return;
    }
    double t[1] = {part};
    block_partials<1>(t, partials);
}

template <typename T, int MODE>
static void sell3_go(hipStream_t st, bool nt, int nblk, int nthr, const Sell3& A, const T* val, const T* x, T* y,
                     const T* f, double* partials, Scalars* sc) {
    if (nt)
        hipLaunchKernelGGL((k_sell3<T, MODE, true>), dim3(nblk), dim3(nthr), 0, st, A.n_rows, A.n_slices, A.sptr,
                           A.scol, val, x, y, f, partials, sc);
    else
        hipLaunchKernelGGL((k_sell3<T, MODE, false>), dim3(nblk), dim3(nthr), 0, st, A.n_rows, A.n_slices, A.sptr,
                           A.scol, val, x, y, f, partials, sc);
}

template <typename T>
void launch_sell3(hipStream_t st, int mode, const Sell3& A, const T* val, const T* x, T* y, const T* f,
                  double* partials, Scalars* sc) {
    const int per = A.waves_per_block;
    const int nblk = (A.n_slices + per - 1) / per;
    const int nthr = per * kWave;
    const bool nt = (int64_t)sizeof(T) * 9 * kWave * A.n_slots > kSellStreamBytes;
    switch (mode) {
        case 0:
            sell3_go<T, 0>(st, nt, nblk, nthr, A, val, x, y, f, partials, sc);
            break;
        case 1:
            sell3_go<T, 1>(st, nt, nblk, nthr, A, val, x, y, f, partials, sc);
            finalize(st, partials, nblk, FIN_PQ, sc);
            break;
        case 2:
            sell3_go<T, 2>(st, nt, nblk, nthr, A, val, x, y, f, partials, sc);
            finalize(st, partials, nblk, FIN_RR, sc);
            break;
        case 4:
            sell3_go<T, 4>(st, nt, nblk, nthr, A, val, x, y, f, partials, sc);
            finalize(st, partials, nblk, FIN_CG1, sc);
            break;
        case 5:
            sell3_go<T, 5>(st, nt, nblk, nthr, A, val, x, y, f, partials, sc);
            finalize(st, partials, nblk, FIN_RU_RES, sc);
            break;
        case 6:
            sell3_go<T, 6>(st, nt, nblk, nthr, A, val, x, y, f, partials, sc);
            finalize(st, partials, nblk, FIN_CG1_RU, sc);
            break;
        case 7:
            sell3_go<T, 7>(st, nt, nblk, nthr, A, val, x, y, f, partials, sc);
            finalize(st, partials, nblk, FIN_PQR, sc);
            break;
        default:
            sell3_go<T, 3>(st, nt, nblk, nthr, A, val, x, y, f, partials, sc);
            break;
    }
}

template void launch_sell3<double>(hipStream_t, int, const Sell3&, const double*, const double*, double*,
                                   const double*, double*, Scalars*);
template void launch_sell3<float>(hipStream_t, int, const Sell3&, const float*, const float*, float*, const float*,
                                  double*, Scalars*);

// ---------------------------------------------------------------------------
// FP16 row-scaled copy of the sliced-ELL operator and its SpMV.  The
// multigrid cycle only preconditions the FP32 flexible CG (whose own
// operator products stay FP32, and whose convergence the FP64 residual
// decides), so its smoothing and restriction residuals can read 2-byte
// values: 22 instead of 40 bytes per 3x3 block and lane.  Row scaling by the
// diagonal keeps every entry O(1): |K_dj| / K_dd <= sqrt(K_jj / K_dd) for an
// SPD operator, far inside FP16 range, and penalised rows become a unit
// diagonal with tiny couplings.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kWave) void k_to_f16_rowscaled(int n_rows, const int32_t* __restrict__ sptr,
                                                            const double* __restrict__ K,
                                                            const double* __restrict__ dinv64,
                                                            __half* __restrict__ K16) {
    const int slice = blockIdx.x, lane = threadIdx.x;
    const int r = slice * kWave + lane;
    double s[3] = {0.0, 0.0, 0.0};
    if (r < n_rows)
        for (int a = 0; a < 3; ++a) s[a] = dinv64[3 * (int64_t)r + a];
    for (int k = sptr[slice]; k < sptr[slice + 1]; ++k)
        for (int c = 0; c < 9; ++c) {
            const int64_t i = ((int64_t)9 * k + c) * kWave + lane;
            K16[i] = __float2half((float)(K[i] * s[c / 3]));
        }
}

void launch_to_f16_rowscaled(hipStream_t st, int n_rows, int n_slices, const int32_t* sptr, const double* K,
                             const double* dinv64, __half* K16) {
    hipLaunchKernelGGL(k_to_f16_rowscaled, dim3(n_slices), dim3(kWave), 0, st, n_rows, sptr, K, dinv64, K16);
}

template <int MODE, bool NT>
__global__ __launch_bounds__(kSellMaxWaves * kWave) void k_sell3h(int n_rows, int n_slices,
                                                                  const int32_t* __restrict__ sptr,
                                                                  const int32_t* __restrict__ scol,
                                                                  const __half* __restrict__ val,
                                                                  const float* __restrict__ dinv,
                                                                  const float* __restrict__ x, float* __restrict__ y,
                                                                  const float* __restrict__ f) {
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int slice = lb * (int)(blockDim.x / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int r = slice * kWave + lane;
    if (slice >= n_slices) return;
    const int k0 = sptr[slice], k1 = sptr[slice + 1];
    const int32_t* cp = scol + (int64_t)k0 * kWave + lane;
    const __half* vp = val + (int64_t)k0 * 9 * kWave + lane;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll 2
    for (int k = k0; k < k1; ++k) {
        const int c = sell_load<NT>(cp);
        const float x0 = x[3 * (int64_t)c], x1 = x[3 * (int64_t)c + 1], x2 = x[3 * (int64_t)c + 2];
        float v[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const unsigned short h = sell_load<NT>((const unsigned short*)(vp + q * kWave));
            v[q] = __half2float(__ushort_as_half(h));
        }
        a0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
        a1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
        a2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
        cp += kWave;
        vp += 9 * kWave;
    }
    if (r >= n_rows) return;
    const int64_t d = 3 * (int64_t)r;
    a0 /= dinv[d];
    a1 /= dinv[d + 1];
    a2 /= dinv[d + 2];
    if (MODE == 3) {
        y[d] = f[d] - a0;
        y[d + 1] = f[d + 1] - a1;
        y[d + 2] = f[d + 2] - a2;
    } else {
        y[d] = a0;
        y[d + 1] = a1;
        y[d + 2] = a2;
    }
}

// Timing of the FP16 cycle SpMV with the residual epilogue (mode 3, the
// most frequent kernel of the multigrid solve) inside a whole pipeline: while
// enabled, each launch outside stream capture is bracketed by two events on
// its own stream; kernel_timer_stop sums their elapsed times.
namespace {
struct CycleTimer {
    bool on = false;
    std::vector<hipEvent_t> ev;  // pairs
    size_t used = 0;
} g_cycle_timer;
}  // namespace

void kernel_timer_start() {
    g_cycle_timer.on = true;
    g_cycle_timer.used = 0;
}

int kernel_timer_stop(int64_t* launches, double* total_ms) {
    CycleTimer& t = g_cycle_timer;
    t.on = false;
    double ms = 0.0;
    for (size_t i = 0; i + 1 < t.used; i += 2) {
        if (hipEventSynchronize(t.ev[i + 1]) != hipSuccess) return -1;
        float e = 0.0f;
        if (hipEventElapsedTime(&e, t.ev[i], t.ev[i + 1]) != hipSuccess) return -1;
        ms += e;
    }
    *launches = (int64_t)(t.used / 2);
    *total_ms = ms;
    t.used = 0;
    return 0;
}

static hipEvent_t cycle_timer_event() {
    CycleTimer& t = g_cycle_timer;
    if (t.used == t.ev.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        t.ev.push_back(e);
    }
    return t.ev[t.used++];
}

void launch_sell3_h(hipStream_t st, int mode, const Sell3& A, const __half* val, const float* dinv32, const float* x,
                    float* y, const float* f) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (g_cycle_timer.on && mode == 3) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
            e0 = cycle_timer_event();
            e1 = cycle_timer_event();
            if (e0 && e1) (void)hipEventRecord(e0, st);
        }
    }
    const int per = A.waves_per_block;
    const int nblk = (A.n_slices + per - 1) / per;
    const int nthr = per * kWave;
    const bool nt = (int64_t)sizeof(__half) * 9 * kWave * A.n_slots > kSellStreamBytes;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(nblk), dim3(nthr), 0, st, A.n_rows, A.n_slices, A.sptr, A.scol, val, dinv32, x,
                           y, f);
    };
    if (mode == 3) {
        if (nt) go(k_sell3h<3, true>);
        else go(k_sell3h<3, false>);
    } else {
        if (nt) go(k_sell3h<0, true>);
        else go(k_sell3h<0, false>);
    }
    if (e0 && e1) (void)hipEventRecord(e1, st);
}

// ---------------------------------------------------------------------------
// Block-row SpMV, b = 3, G lanes per block row, component-major values.
// MODE 0: y = K x
// MODE 1: y = K x and (x . y) -> finalize: sc->pq, sc->alpha = sc->rz / pq
// MODE 2: y = f - K x (x = current solution, f = load) -> finalize: sc->rr = |y|^2
// MODE 3: y = f - K x without the norm
// MODE 4 (sliced-ELL only): y = K x with x = u = Dinv r and f = r, partials of
//   r.u, y.u and r.r -> finalize: the single-reduction PCG step scalars
// MODE 5 (sliced-ELL only): MODE 2 when sc->gate is set (mixed solve's
//   reliable update) -> finalize: sc->rr, convergence on the FP64 residual
// MODE 6 (sliced-ELL only): MODE 4 paused while sc->gate is set -> finalize:
//   the mixed solve's FP32 step scalars, or a refresh request
// MODE 7 (sliced-ELL only): y = K x with x = p and f = r, partials of p.y and
//   p.r -> finalize: the flexible PCG step length alpha = p.r / p.q
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

// PCG with a preconditioner applied by its own kernels (z = M^-1 r formed
// between these launches; the LDL^T block preconditioner): x = 0, r = b.
__global__ __launch_bounds__(kBlock) void k_pcgz_start(int64_t n, const double* __restrict__ b,
                                                       double* __restrict__ x, double* __restrict__ r) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] = 0.0;
        r[i] = b[i];
    }
}

// partials of r.z and r.r; at the start also p = z
__global__ __launch_bounds__(kBlock) void k_pcgz_dots(int64_t n, const double* __restrict__ r,
                                                      const double* __restrict__ z, double* __restrict__ p, int init,
                                                      double* partials, const Scalars* sc) {
    if (!init && sc->done) return;
    double t[2] = {0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double ri = r[i], zi = z[i];
        if (init) p[i] = zi;
        t[0] += ri * zi;
        t[1] += ri * ri;
    }
    block_partials<2>(t, partials);
}

// x += alpha p; r -= alpha q
__global__ __launch_bounds__(kBlock) void k_pcgz_update(int64_t n, double* __restrict__ x, double* __restrict__ r,
                                                        const double* __restrict__ p, const double* __restrict__ q,
                                                        const Scalars* __restrict__ sc) {
    if (sc->done) return;
    const double alpha = sc->alpha;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        r[i] -= alpha * q[i];
    }
}

// p = z + beta p
__global__ __launch_bounds__(kBlock) void k_pcgz_dir(int64_t n, const double* __restrict__ z, double* __restrict__ p,
                                                     const Scalars* __restrict__ sc) {
    if (sc->done) return;
    const double beta = sc->beta;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = z[i] + beta * p[i];
}

// Single-reduction (Chronopoulos-Gear) Jacobi PCG: one SpMV with all three
// dot products fused (K u with r.u, (K u).u, r.r) and one vector update per
// iteration, against an SpMV, two vector kernels and two reductions for the
// textbook recurrence.  Same Krylov iterates in exact arithmetic:
//   beta = gamma / gamma_old, alpha = gamma / (delta - beta gamma / alpha_old)
//   p = u + beta p ; s = w + beta s (= K p) ; x += alpha p ; r -= alpha s ;
//   u = Dinv r
// Start: cold (x = 0, r = b) or warm (r = b - K x already formed in r);
// u = Dinv r, p = s = 0; partials of |b|^2.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg1_init(int64_t n, const T* __restrict__ b, int cold,
                                                     const T* __restrict__ dinv, T* __restrict__ x, T* __restrict__ r,
                                                     T* __restrict__ u, T* __restrict__ p, T* __restrict__ s,
                                                     double* partials) {
    double t[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T bi = b[i];
        T ri;
        if (cold) {
            ri = bi;
            r[i] = ri;
            x[i] = 0;
        } else {
            ri = r[i];
        }
        u[i] = dinv[i] * ri;
        p[i] = 0;
        s[i] = 0;
        t[0] += (double)bi * bi;
    }
    block_partials<1>(t, partials);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg1_update(int64_t n, const T* __restrict__ dinv, T* __restrict__ x,
                                                       T* __restrict__ r, T* __restrict__ u, const T* __restrict__ w,
                                                       T* __restrict__ p, T* __restrict__ s,
                                                       const Scalars* __restrict__ sc) {
    if (sc->done) return;
    const T alpha = (T)sc->alpha, beta = (T)sc->beta;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T pi = u[i] + beta * p[i];
        const T si = w[i] + beta * s[i];
        p[i] = pi;
        s[i] = si;
        x[i] += alpha * pi;
        const T ri = r[i] - alpha * si;
        r[i] = ri;
        u[i] = dinv[i] * ri;
    }
}

// Flexible Jacobi PCG (one truncation vector): alpha = p.r / p.q (launch_sell3
// mode 7), x += alpha p, r -= alpha q, beta = -(Dinv r).q / p.q, then
// p = Dinv r + beta p (launch_pcg_dir).  Fresh inner products every
// iteration, so a residual replaced by a reliable update needs no repair.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_fpcg_update(int64_t n, const T* __restrict__ dinv, T* __restrict__ x,
                                                        T* __restrict__ r, const T* __restrict__ p,
                                                        const T* __restrict__ q, double* partials,
                                                        const Scalars* sc) {
    if (sc->done) return;
    const T alpha = (T)sc->alpha;
    double t[2] = {0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        const T ri = r[i] - alpha * q[i];
        r[i] = ri;
        t[0] += (double)(dinv[i] * ri) * q[i];
        t[1] += (double)ri * ri;
    }
    block_partials<2>(t, partials);
}
// reliable update: r = b (the reloaded residual); p *= ratio, or p = Dinv r
// when ratio = 0 (restart)
template <typename T>
__global__ void k_ru_reload(int64_t n, const T* __restrict__ b, const T* __restrict__ dinv, T* __restrict__ r,
                            T* __restrict__ p, double ratio) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T bi = b[i];
        r[i] = bi;
        p[i] = ratio != 0.0 ? (T)(p[i] * ratio) : dinv[i] * bi;
    }
}

template <typename T>
void launch_fpcg_update(hipStream_t st, int64_t n, const T* dinv, T* x, T* r, const T* p, const T* q,
                        double* partials, Scalars* sc) {
    hipLaunchKernelGGL(k_fpcg_update<T>, dim3(kVecBlocks), dim3(kBlock), 0, st, n, dinv, x, r, p, q, partials, sc);
    finalize(st, partials, kVecBlocks, FIN_FUPD, sc);
}
template <typename T>
void launch_ru_reload(hipStream_t st, int64_t n, const T* b, const T* dinv, T* r, T* p, double ratio) {
    hipLaunchKernelGGL(k_ru_reload<T>, dim3(kVecBlocks), dim3(kBlock), 0, st, n, b, dinv, r, p, ratio);
}

template <typename T>
void launch_cg1_init(hipStream_t st, int64_t n, const T* b, bool cold, const T* dinv, T* x, T* r, T* u, T* p, T* s,
                     double* partials, Scalars* sc) {
    hipLaunchKernelGGL(k_cg1_init<T>, dim3(kVecBlocks), dim3(kBlock), 0, st, n, b, cold ? 1 : 0, dinv, x, r, u, p, s,
                       partials);
    finalize(st, partials, kVecBlocks, FIN_CG1_INIT, sc);
}

// --- mixed-precision solve under device-side control ----------------------
__global__ void k_ru_start(Scalars* __restrict__ sc, double ru2) {
    sc->ff = sc->dot;
    sc->rr = sc->dot;
    sc->zq = ru2;
    sc->it = 0;
    sc->n_resid = 0;
    sc->rz = 0.0;
    sc->done = sc->dot == 0.0;
    sc->gate = !sc->done;
}
// FP32 step with the solution in FP64: p = u + beta p ; s = w + beta s ;
// x += alpha p ; r -= alpha s ; u = Dinv r
__global__ __launch_bounds__(kBlock) void k_cg1_update_mx(int64_t n, const float* __restrict__ dinv,
                                                          double* __restrict__ x, float* __restrict__ r,
                                                          float* __restrict__ u, const float* __restrict__ w,
                                                          float* __restrict__ p, float* __restrict__ s,
                                                          const Scalars* __restrict__ sc) {
    if (sc->done || sc->gate) return;
    const float alpha = (float)sc->alpha, beta = (float)sc->beta;
    const double alpha64 = sc->alpha;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float pi = u[i] + beta * p[i];
        const float si = w[i] + beta * s[i];
        p[i] = pi;
        s[i] = si;
        x[i] += alpha64 * (double)pi;
        const float ri = r[i] - alpha * si;
        r[i] = ri;
        u[i] = dinv[i] * ri;
    }
}
// reliable update: r = (float) r64, u = Dinv r (and p = s = 0 at the start)
__global__ __launch_bounds__(kBlock) void k_ru_load(int64_t n, const double* __restrict__ r64,
                                                    const float* __restrict__ dinv, float* __restrict__ r,
                                                    float* __restrict__ u, float* __restrict__ p,
                                                    float* __restrict__ s, int restart,
                                                    const Scalars* __restrict__ sc) {
    if (!sc->gate) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float ri = (float)r64[i];
        r[i] = ri;
        u[i] = dinv[i] * ri;
        if (restart) {
            p[i] = 0.f;
            s[i] = 0.f;
        }
    }
}
__global__ void k_ru_clear(Scalars* __restrict__ sc) { sc->gate = 0; }

static void mixed_refresh(hipStream_t st, const MixedSolve& m, int restart) {
    launch_sell3<double>(st, 5, *m.A, m.K64, m.x64, m.r64, m.rhs, m.partials, m.sc);
    hipLaunchKernelGGL(k_ru_load, dim3(kVecBlocks), dim3(kBlock), 0, st, m.n, m.r64, m.dinv32, m.r32, m.u32, m.p32,
                       m.s32, restart, (const Scalars*)m.sc);
    hipLaunchKernelGGL(k_ru_clear, dim3(1), dim3(1), 0, st, m.sc);
}

void launch_mixed_start(hipStream_t st, const MixedSolve& m) {
    launch_dot64(st, m.n, m.rhs, m.rhs, m.partials, m.sc);
    hipLaunchKernelGGL(k_ru_start, dim3(1), dim3(1), 0, st, m.sc, m.ru_factor * m.ru_factor);
    mixed_refresh(st, m, 1);
}

void launch_mixed_iteration(hipStream_t st, const MixedSolve& m) {
    launch_sell3<float>(st, 6, *m.A, m.K32, m.u32, m.w32, m.r32, m.partials, m.sc);
    hipLaunchKernelGGL(k_cg1_update_mx, dim3(kVecBlocks), dim3(kBlock), 0, st, m.n, m.dinv32, m.x64, m.r32, m.u32,
                       m.w32, m.p32, m.s32, (const Scalars*)m.sc);
}

void launch_mixed_refresh(hipStream_t st, const MixedSolve& m) { mixed_refresh(st, m, 0); }
template <typename T>
void launch_cg1_update(hipStream_t st, int64_t n, const T* dinv, T* x, T* r, T* u, const T* w, T* p, T* s,
                       const Scalars* sc) {
    hipLaunchKernelGGL(k_cg1_update<T>, dim3(kVecBlocks), dim3(kBlock), 0, st, n, dinv, x, r, u, w, p, s, sc);
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

void launch_pcgz_start(hipStream_t st, int64_t n, const double* b, double* x, double* r) {
    hipLaunchKernelGGL(k_pcgz_start, dim3(kVecBlocks), dim3(kBlock), 0, st, n, b, x, r);
}
void launch_pcgz_dots(hipStream_t st, int64_t n, const double* r, const double* z, double* p, bool init,
                      double* partials, Scalars* sc) {
    hipLaunchKernelGGL(k_pcgz_dots, dim3(kVecBlocks), dim3(kBlock), 0, st, n, r, z, p, init ? 1 : 0, partials, sc);
    finalize(st, partials, kVecBlocks, init ? FIN_INIT : FIN_UPDATE, sc);
}
void launch_pcgz_update(hipStream_t st, int64_t n, double* x, double* r, const double* p, const double* q,
                        const Scalars* sc) {
    hipLaunchKernelGGL(k_pcgz_update, dim3(kVecBlocks), dim3(kBlock), 0, st, n, x, r, p, q, sc);
}
void launch_pcgz_dir(hipStream_t st, int64_t n, const double* z, double* p, const Scalars* sc) {
    hipLaunchKernelGGL(k_pcgz_dir, dim3(kVecBlocks), dim3(kBlock), 0, st, n, z, p, sc);
}

#define INST_PCG(T)                                                                                               \
    template void launch_pcg_init<T>(hipStream_t, int64_t, const T*, const T*, T*, T*, T*, double*, Scalars*);    \
    template void launch_pcg_update<T>(hipStream_t, int64_t, const T*, T*, T*, const T*, const T*, double*,       \
                                       Scalars*);                                                                 \
    template void launch_pcg_dir<T>(hipStream_t, int64_t, const T*, const T*, T*, const Scalars*);               \
    template void launch_cg1_init<T>(hipStream_t, int64_t, const T*, bool, const T*, T*, T*, T*, T*, T*, double*, \
                                     Scalars*);                                                                   \
    template void launch_cg1_update<T>(hipStream_t, int64_t, const T*, T*, T*, T*, const T*, T*, T*,              \
                                       const Scalars*);                                                           \
    template void launch_fpcg_update<T>(hipStream_t, int64_t, const T*, T*, T*, const T*, const T*, double*,      \
                                        Scalars*);                                                                \
    template void launch_ru_reload<T>(hipStream_t, int64_t, const T*, const T*, T*, T*, double);
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

// Material-parameter sensitivities of -lambda^T K_e u for every brick8
// element, at its isotropic (E, nu, G = E/(2(1+nu))) point, with respect to
// the nine parameters of the orthotropic compliance S the reference builds
// (E1, E2, E3, nu12, nu13, nu23, G12, G13, G23; S/element_properties.rs:
// 728-759): since dC/dp = -C (dS/dp) C, each is
// sum_gp detJ sigma(lambda)^T (dS/dp) sigma(u) with sigma = C eps (Voigt
// [11, 22, 33, 12, 13, 23], engineering shears).  out[9 e + p].
__global__ __launch_bounds__(128) void k_hex8_sens_mat(int n_el, const int32_t* __restrict__ conn_r,
                                                       const double* __restrict__ xyz_r,
                                                       const double* __restrict__ E, const double* __restrict__ nu,
                                                       const double* __restrict__ u, const double* __restrict__ lam,
                                                       double* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_el) return;
    double X[8][3], U[8][3], L[8][3];
    for (int a = 0; a < 8; ++a) {
        const int64_t r = conn_r[8 * (int64_t)e + a];
        for (int c = 0; c < 3; ++c) {
            X[a][c] = xyz_r[3 * r + c];
            U[a][c] = u[3 * r + c];
            L[a][c] = lam[3 * r + c];
        }
    }
    const double Ee = E[e], ne = nu[e];
    const double lm = Ee * ne / ((1.0 + ne) * (1.0 - 2.0 * ne));
    const double G = Ee / (2.0 * (1.0 + ne));
    const double iE = 1.0 / Ee, iE2 = iE * iE, iG2 = 1.0 / (G * G);
    const double g = 0.5773502691896257;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int gp = 0; gp < 8; ++gp) {
        const double s0 = (gp & 1) ? g : -g, s1 = (gp & 2) ? g : -g, s2 = (gp & 4) ? g : -g;
        double dN[8][3];
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int a = 0; a < 8; ++a) {
            const double fx = 1.0 + kHexNdSign[a][0] * s0, fy = 1.0 + kHexNdSign[a][1] * s1,
                         fz = 1.0 + kHexNdSign[a][2] * s2;
            dN[a][0] = 0.125 * kHexNdSign[a][0] * fy * fz;
            dN[a][1] = 0.125 * fx * kHexNdSign[a][1] * fz;
            dN[a][2] = 0.125 * fx * fy * kHexNdSign[a][2];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) J[i][j] += X[a][i] * dN[a][j];
        }
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        const double id = 1.0 / det;
        const double Ji[3][3] = {{c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id,
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
        // stresses of u and lambda
        const double tu = gu[0][0] + gu[1][1] + gu[2][2], tl = gl[0][0] + gl[1][1] + gl[2][2];
        const double su[6] = {lm * tu + 2.0 * G * gu[0][0], lm * tu + 2.0 * G * gu[1][1], lm * tu + 2.0 * G * gu[2][2],
                              G * (gu[0][1] + gu[1][0]), G * (gu[0][2] + gu[2][0]), G * (gu[1][2] + gu[2][1])};
        const double sl[6] = {lm * tl + 2.0 * G * gl[0][0], lm * tl + 2.0 * G * gl[1][1], lm * tl + 2.0 * G * gl[2][2],
                              G * (gl[0][1] + gl[1][0]), G * (gl[0][2] + gl[2][0]), G * (gl[1][2] + gl[2][1])};
        const double p01 = sl[0] * su[1] + sl[1] * su[0], p02 = sl[0] * su[2] + sl[2] * su[0],
                     p12 = sl[1] * su[2] + sl[2] * su[1];
        // S00 = 1/E1, S01 = -nu12/E1, S02 = -nu13/E1, S11 = 1/E2, S12 = -nu23/E2,
        // S22 = 1/E3, S33 = 1/G12, S44 = 1/G13, S55 = 1/G23 (symmetric)
        acc[0] += det * (-iE2 * sl[0] * su[0] + ne * iE2 * (p01 + p02));
        acc[1] += det * (-iE2 * sl[1] * su[1] + ne * iE2 * p12);
        acc[2] += det * (-iE2 * sl[2] * su[2]);
        acc[3] += det * (-iE * p01);
        acc[4] += det * (-iE * p02);
        acc[5] += det * (-iE * p12);
        acc[6] += det * (-iG2 * sl[3] * su[3]);
        acc[7] += det * (-iG2 * sl[4] * su[4]);
        acc[8] += det * (-iG2 * sl[5] * su[5]);
    }
    for (int p = 0; p < 9; ++p) out[9 * (int64_t)e + p] = acc[p];
}

// Element fields at the element centre (natural coordinates 0, 0, 0: the
// brick8 stress point s_cent of get_stress_strain_dfd0,
// S/element_soln_fields.rs:1463-1530, S/element_meth.rs:81-83), one thread per
// element: gradients g_t = J^-T dN_t at the centre, strain
// (e11, e22, e33, g12, g13, g23) with engineering shear
// (get_solid_strain_dfd0, :1345-1461), stress = C strain for the element's
// isotropic C, and the element volume sum_gp detJ (get_volume_dfd0).
// out[13 e ..]: 6 strains, 6 stresses, volume.
__device__ __forceinline__ void hex8_centre_grad(const double (&X)[8][3], double (&g)[8][3]) {
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int t = 0; t < 8; ++t) {
        double sx, sy, sz;
        hex_sign(t, sx, sy, sz);
        const double dn[3] = {0.125 * sx, 0.125 * sy, 0.125 * sz};
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 3; ++c) J[3 * i + c] += X[t][i] * dn[c];
    }
    const double c00 = J[4] * J[8] - J[5] * J[7], c01 = J[5] * J[6] - J[3] * J[8], c02 = J[3] * J[7] - J[4] * J[6];
    const double id = 1.0 / (J[0] * c00 + J[1] * c01 + J[2] * c02);
    const double Ji[9] = {c00 * id, (J[2] * J[7] - J[1] * J[8]) * id, (J[1] * J[5] - J[2] * J[4]) * id,
                          c01 * id, (J[0] * J[8] - J[2] * J[6]) * id, (J[2] * J[3] - J[0] * J[5]) * id,
                          c02 * id, (J[1] * J[6] - J[0] * J[7]) * id, (J[0] * J[4] - J[1] * J[3]) * id};
    for (int t = 0; t < 8; ++t) {
        double sx, sy, sz;
        hex_sign(t, sx, sy, sz);
        const double dn[3] = {0.125 * sx, 0.125 * sy, 0.125 * sz};
        for (int j = 0; j < 3; ++j) g[t][j] = dn[0] * Ji[j] + dn[1] * Ji[3 + j] + dn[2] * Ji[6 + j];
    }
}

__device__ __forceinline__ void hex8_load_x(const int32_t* __restrict__ conn_r, const double* __restrict__ xyz_r,
                                            int e, double (&X)[8][3]) {
    for (int t = 0; t < 8; ++t) {
        const int node = conn_r[8 * (int64_t)e + t];
        for (int c = 0; c < 3; ++c) X[t][c] = xyz_r[3 * (int64_t)node + c];
    }
}

__global__ __launch_bounds__(256) void k_hex8_centre_fields(int n_el, const int32_t* __restrict__ conn_r,
                                                            const double* __restrict__ xyz_r,
                                                            const double* __restrict__ E,
                                                            const double* __restrict__ nu,
                                                            const double* __restrict__ u, double* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_el) return;
    double X[8][3], g[8][3];
    hex8_load_x(conn_r, xyz_r, e, X);
    hex8_centre_grad(X, g);
    double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // du_i/dx_j
    for (int t = 0; t < 8; ++t) {
        const int node = conn_r[8 * (int64_t)e + t];
        for (int i = 0; i < 3; ++i) {
            const double ui = u[3 * (int64_t)node + i];
            for (int j = 0; j < 3; ++j) G[3 * i + j] += ui * g[t][j];
        }
    }
    const double eps[6] = {G[0], G[4], G[8], G[1] + G[3], G[2] + G[6], G[5] + G[7]};
    const double Ee = E[e], ne = nu[e];
    const double lam = Ee * ne / ((1.0 + ne) * (1.0 - 2.0 * ne)), mu = Ee / (2.0 * (1.0 + ne));
    const double tr = eps[0] + eps[1] + eps[2];
    double* o = out + 13 * (int64_t)e;
    for (int k = 0; k < 6; ++k) o[k] = eps[k];
    for (int k = 0; k < 3; ++k) o[6 + k] = lam * tr + 2.0 * mu * eps[k];
    for (int k = 3; k < 6; ++k) o[6 + k] = mu * eps[k];
    // volume: detJ at the 8 Gauss points (weights 1)
    const double gq = 0.5773502691896257;
    double vol = 0.0;
    for (int gp = 0; gp < 8; ++gp) {
        const double p0 = (gp & 1) ? gq : -gq, p1 = (gp & 2) ? gq : -gq, p2 = (gp & 4) ? gq : -gq;
        double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int t = 0; t < 8; ++t) {
            double sx, sy, sz;
            hex_sign(t, sx, sy, sz);
            const double fx = 1.0 + sx * p0, fy = 1.0 + sy * p1, fz = 1.0 + sz * p2;
            const double dn[3] = {0.125 * sx * fy * fz, 0.125 * fx * sy * fz, 0.125 * fx * fy * sz};
            for (int i = 0; i < 3; ++i)
                for (int c = 0; c < 3; ++c) J[3 * i + c] += X[t][i] * dn[c];
        }
        vol += J[0] * (J[4] * J[8] - J[5] * J[7]) + J[1] * (J[5] * J[6] - J[3] * J[8]) +
               J[2] * (J[3] * J[7] - J[4] * J[6]);
    }
    o[12] = vol;
}

// Adjoint load of centre-point element quantities: r_e = B_c^T h_e for the
// element's strain-space weights h_e (6 per element), B_c the centre strain
// operator of hex8_centre_grad, written as 24 values per element (node t,
// component i at 3 t + i).
__global__ __launch_bounds__(256) void k_hex8_centre_adjoint(int n_el, const int32_t* __restrict__ conn_r,
                                                             const double* __restrict__ xyz_r,
                                                             const double* __restrict__ h,
                                                             double* __restrict__ re) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_el) return;
    double X[8][3], g[8][3];
    hex8_load_x(conn_r, xyz_r, e, X);
    hex8_centre_grad(X, g);
    const double* w = h + 6 * (int64_t)e;
    double* r = re + 24 * (int64_t)e;
    for (int t = 0; t < 8; ++t) {
        r[3 * t + 0] = w[0] * g[t][0] + w[3] * g[t][1] + w[4] * g[t][2];
        r[3 * t + 1] = w[1] * g[t][1] + w[3] * g[t][0] + w[5] * g[t][2];
        r[3 * t + 2] = w[2] * g[t][2] + w[4] * g[t][0] + w[5] * g[t][1];
    }
}

// Node-row gather of per-element 24-vectors: y[3 r + i] = sum over the
// records 8 e + a of row r of re[24 e + 3 a + i], in record order.
__global__ __launch_bounds__(256) void k_elem_vec_gather(int n_rows, const int32_t* __restrict__ n2e_ptr,
                                                         const int32_t* __restrict__ n2rec,
                                                         const double* __restrict__ re, double* __restrict__ y) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int k = n2e_ptr[r]; k < n2e_ptr[r + 1]; ++k) {
        const int64_t rec = n2rec[k];
        const int64_t e = rec >> 3, a = rec & 7;
        for (int i = 0; i < 3; ++i) acc[i] += re[24 * e + 3 * a + i];
    }
    for (int i = 0; i < 3; ++i) y[3 * (int64_t)r + i] = acc[i];
}

// Shape sensitivities of the brick8 element energy form: for element e and
// node coordinate X_{a,c} (direction d = 3 a + c, one thread per (e, d)),
// out[24 e + d] = lam_e^T (dK_e/dX_{a,c}) u_e, by forward-mode dual numbers
// through the Jacobian, its inverse, the physical gradients and detJ at the 8
// Gauss points of K_e = sum_gp detJ B^T C B (isotropic C of E_e, nu_e): the
// d_relasticd_d coordinate duals of nodeCoord design variables
// (S/model_analysis.rs:1379-1457, Node::get_crd_dfd0 S/node_meth.rs:255-282).
struct Dn {
    double v, d;
};
__device__ __forceinline__ Dn dn_add(Dn a, Dn b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dn dn_sub(Dn a, Dn b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dn dn_mul(Dn a, Dn b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dn dn_scale(Dn a, double s) { return {a.v * s, a.d * s}; }

__global__ __launch_bounds__(256) void k_hex8_shape_sens(int n_el, const int32_t* __restrict__ conn_r,
                                                         const double* __restrict__ xyz_r,
                                                         const double* __restrict__ E, const double* __restrict__ nu,
                                                         const double* __restrict__ u, const double* __restrict__ lam,
                                                         double* __restrict__ out) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= 24 * (int64_t)n_el) return;
    const int e = (int)(tid / 24), dir = (int)(tid % 24);
    const int sa = dir / 3, sc = dir % 3;  // seeded node and coordinate
    double X[8][3], U[8][3], Lm[8][3];
    for (int t = 0; t < 8; ++t) {
        const int node = conn_r[8 * (int64_t)e + t];
        for (int c = 0; c < 3; ++c) {
            X[t][c] = xyz_r[3 * (int64_t)node + c];
            U[t][c] = u[3 * (int64_t)node + c];
            Lm[t][c] = lam[3 * (int64_t)node + c];
        }
    }
    const double Ee = E[e], ne = nu[e];
    const double lm = Ee * ne / ((1.0 + ne) * (1.0 - 2.0 * ne)), mu = Ee / (2.0 * (1.0 + ne));
    const double g = 0.5773502691896257;
    double acc = 0.0;
    for (int gp = 0; gp < 8; ++gp) {
        const double p0 = (gp & 1) ? g : -g, p1 = (gp & 2) ? g : -g, p2 = (gp & 4) ? g : -g;
        double dn[8][3];
        for (int t = 0; t < 8; ++t) {
            double sx, sy, sz;
            hex_sign(t, sx, sy, sz);
            const double fx = 1.0 + sx * p0, fy = 1.0 + sy * p1, fz = 1.0 + sz * p2;
            dn[t][0] = 0.125 * sx * fy * fz;
            dn[t][1] = 0.125 * fx * sy * fz;
            dn[t][2] = 0.125 * fx * fy * sz;
        }
        // J[i][c] = sum_t X_t[i] dn_t[c]; its derivative along X_{sa,sc} is
        // dn_sa[c] in row sc
        Dn J[9];
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 3; ++c) {
                double v = 0.0;
                for (int t = 0; t < 8; ++t) v += X[t][i] * dn[t][c];
                J[3 * i + c] = {v, i == sc ? dn[sa][c] : 0.0};
            }
        const Dn c00 = dn_sub(dn_mul(J[4], J[8]), dn_mul(J[5], J[7]));
        const Dn c01 = dn_sub(dn_mul(J[5], J[6]), dn_mul(J[3], J[8]));
        const Dn c02 = dn_sub(dn_mul(J[3], J[7]), dn_mul(J[4], J[6]));
        const Dn det = dn_add(dn_add(dn_mul(J[0], c00), dn_mul(J[1], c01)), dn_mul(J[2], c02));
        const Dn id = {1.0 / det.v, -det.d / (det.v * det.v)};
        const Dn Ji[9] = {dn_mul(c00, id),
                          dn_mul(dn_sub(dn_mul(J[2], J[7]), dn_mul(J[1], J[8])), id),
                          dn_mul(dn_sub(dn_mul(J[1], J[5]), dn_mul(J[2], J[4])), id),
                          dn_mul(c01, id),
                          dn_mul(dn_sub(dn_mul(J[0], J[8]), dn_mul(J[2], J[6])), id),
                          dn_mul(dn_sub(dn_mul(J[2], J[3]), dn_mul(J[0], J[5])), id),
                          dn_mul(c02, id),
                          dn_mul(dn_sub(dn_mul(J[1], J[6]), dn_mul(J[0], J[7])), id),
                          dn_mul(dn_sub(dn_mul(J[0], J[4]), dn_mul(J[1], J[3])), id)};
        // displacement gradients of u and lam: sum_t U_t[i] g_t[j],
        // g_t[j] = sum_c dn_t[c] Ji[c][j]
        Dn Gu[9], Gl[9];
        for (int k = 0; k < 9; ++k) Gu[k] = Gl[k] = {0.0, 0.0};
        for (int t = 0; t < 8; ++t) {
            Dn gt[3];
            for (int j = 0; j < 3; ++j)
                gt[j] = dn_add(dn_add(dn_scale(Ji[j], dn[t][0]), dn_scale(Ji[3 + j], dn[t][1])),
                               dn_scale(Ji[6 + j], dn[t][2]));
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    Gu[3 * i + j] = dn_add(Gu[3 * i + j], dn_scale(gt[j], U[t][i]));
                    Gl[3 * i + j] = dn_add(Gl[3 * i + j], dn_scale(gt[j], Lm[t][i]));
                }
        }
        const Dn eu[6] = {Gu[0], Gu[4], Gu[8], dn_add(Gu[1], Gu[3]), dn_add(Gu[2], Gu[6]), dn_add(Gu[5], Gu[7])};
        const Dn el[6] = {Gl[0], Gl[4], Gl[8], dn_add(Gl[1], Gl[3]), dn_add(Gl[2], Gl[6]), dn_add(Gl[5], Gl[7])};
        const Dn tr = dn_add(dn_add(eu[0], eu[1]), eu[2]);
        Dn w = {0.0, 0.0};
        for (int k = 0; k < 3; ++k) w = dn_add(w, dn_mul(el[k], dn_add(dn_scale(tr, lm), dn_scale(eu[k], 2.0 * mu))));
        for (int k = 3; k < 6; ++k) w = dn_add(w, dn_mul(el[k], dn_scale(eu[k], mu)));
        acc += det.d * w.v + det.v * w.d;
    }
    out[tid] = acc;
}

void launch_hex8_shape_sens(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                            const double* nu, const double* u, const double* lam, double* out) {
    if (n_el <= 0) return;
    const int64_t n = 24 * (int64_t)n_el;
    hipLaunchKernelGGL(k_hex8_shape_sens, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n_el, conn_r, xyz_r, E,
                       nu, u, lam, out);
}

void launch_hex8_centre_fields(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                               const double* nu, const double* u, double* out) {
    if (n_el <= 0) return;
    hipLaunchKernelGGL(k_hex8_centre_fields, dim3((n_el + 255) / 256), dim3(256), 0, st, n_el, conn_r, xyz_r, E, nu,
                       u, out);
}

void launch_hex8_centre_adjoint(hipStream_t st, int n_el, int n_rows, const int32_t* conn_r, const double* xyz_r,
                                const int32_t* n2e_ptr, const int32_t* n2rec, const double* h, double* re,
                                double* y) {
    if (n_el > 0)
        hipLaunchKernelGGL(k_hex8_centre_adjoint, dim3((n_el + 255) / 256), dim3(256), 0, st, n_el, conn_r, xyz_r, h,
                           re);
    hipLaunchKernelGGL(k_elem_vec_gather, dim3((n_rows + 255) / 256), dim3(256), 0, st, n_rows, n2e_ptr, n2rec, re,
                       y);
}

void launch_hex8_sens_mat(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                          const double* nu, const double* u, const double* lam, double* out) {
    hipLaunchKernelGGL(k_hex8_sens_mat, dim3((n_el + 127) / 128), dim3(128), 0, st, n_el, conn_r, xyz_r, E, nu, u,
                       lam, out);
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
