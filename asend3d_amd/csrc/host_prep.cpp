// Host-side model preparation for the MI355X solve path: node reordering and
// DOF numbering, surface-traction loads, seeded moduli, block size rule.
// These run once per model on the CPU (graph/ordering work with a sequential
// dependency chain); everything per-iteration runs on the GPU.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "asend_gpu.h"
#include "host_prep.hpp"

namespace asend {

// Reference node neighbour lists are append-if-absent linked lists built in
// element order (S/model_prep.rs:76-101, S/nd_el_set_meth.rs:5-14).  The
// insertion order matters for the level-order traversal, so each list keeps
// it; membership is tested with a per-list linear scan exactly like the
// reference (lists are short: <= 26 neighbours for hex meshes).
struct NbrLists {
    std::vector<std::vector<int32_t>> v;
    explicit NbrLists(int64_t n) : v(n) {}
    void add(int32_t a, int32_t b) {
        auto& l = v[a];
        for (int32_t x : l)
            if (x == b) return;
        l.push_back(b);
    }
};

int reorder_nodes(int64_t n_nodes, const double* xyz, int64_t n_el, int nn, int dof_per_nd, const int32_t* conn,
                  int64_t n_const, const int64_t* cterm_ptr, const int64_t* term_set_ptr, const int32_t* set_nodes,
                  int64_t block_dim, int32_t* ordered, int64_t* dof_index, int64_t* n_dof_out) {
    if (n_nodes <= 0 || nn <= 0 || !xyz || !conn || !ordered || !dof_index || block_dim <= 0) return ASEND_ERR_ARG;
    NbrLists nc(n_nodes);
    std::vector<uint8_t> ndof(n_nodes, 3);
    for (int64_t e = 0; e < n_el; ++e) {
        const int32_t* c = conn + (int64_t)nn * e;
        for (int i = 0; i < nn; ++i) {
            if (c[i] < 0 || c[i] >= n_nodes) return ASEND_ERR_ARG;
            if (dof_per_nd > 3) ndof[c[i]] = (uint8_t)dof_per_nd;
            for (int j = i + 1; j < nn; ++j) {
                nc.add(c[i], c[j]);
                nc.add(c[j], c[i]);
            }
        }
    }
    // constraint couplings, S/model_prep.rs:20-54
    for (int64_t ci = 0; ci < n_const; ++ci) {
        for (int64_t t1 = cterm_ptr[ci]; t1 < cterm_ptr[ci + 1]; ++t1) {
            const int32_t* s1 = set_nodes + term_set_ptr[t1];
            int64_t l1 = term_set_ptr[t1 + 1] - term_set_ptr[t1];
            for (int64_t t2 = cterm_ptr[ci]; t2 < cterm_ptr[ci + 1]; ++t2) {
                const int32_t* s2 = set_nodes + term_set_ptr[t2];
                int64_t l2 = term_set_ptr[t2 + 1] - term_set_ptr[t2];
                if (l1 > 1 && l2 > 1) {
                    if (l1 > l2) return ASEND_ERR_ARG;  // mismatched set sizes
                    for (int64_t k = 0; k < l1; ++k) {
                        nc.add(s1[k], s2[k]);
                        nc.add(s2[k], s1[k]);
                    }
                } else {
                    for (int64_t i = 0; i < l1; ++i)
                        for (int64_t j = 0; j < l2; ++j) {
                            nc.add(s1[i], s2[j]);
                            nc.add(s2[j], s1[i]);
                        }
                }
            }
        }
    }
    int64_t min_ct = n_nodes, start = 0;
    for (int64_t i = 0; i < n_nodes; ++i)
        if ((int64_t)nc.v[i].size() < min_ct) {
            min_ct = (int64_t)nc.v[i].size();
            start = i;
        }

    // Uninserted nodes are kept in a compact list so each restart scans only
    // the nodes still outside the ordering.  The scan visits them in
    // ascending label order and keeps the first strict minimum of the same
    // FP64 distance, so ties resolve to the lowest label exactly as the
    // reference's full 0..n scan does (S/model_prep.rs:170-181).
    std::vector<int32_t> remaining(n_nodes);
    std::vector<int32_t> pos(n_nodes);
    for (int64_t i = 0; i < n_nodes; ++i) {
        remaining[i] = (int32_t)i;
        pos[i] = (int32_t)i;
    }
    int64_t n_rem = n_nodes;
    int64_t dead = 0;  // removed entries marked -1, compacted lazily
    std::vector<uint8_t> ins(n_nodes, 0);
    auto insert = [&](int64_t nd) {
        ins[nd] = 1;
        remaining[pos[nd]] = -1;
        ++dead;
        --n_rem;
    };
    auto compact = [&]() {
        int64_t w = 0;
        int64_t len = n_rem + dead;
        for (int64_t r = 0; r < len; ++r) {
            int32_t v = remaining[r];
            if (v >= 0) {
                remaining[w] = v;
                pos[v] = (int32_t)w;
                ++w;
            }
        }
        dead = 0;
    };

    ordered[0] = (int32_t)start;
    insert(start);
    int64_t on_len = 1, on_i = 0, since = 0, nd1 = start, nd2 = start;
    while (on_len < n_nodes) {
        const auto& nb = nc.v[nd1];
        size_t it = 0;
        bool end = false;
        while (!end && since < block_dim) {
            if (it < nb.size()) nd2 = nb[it++];
            else end = true;
            if (!ins[nd2]) {
                ordered[on_len++] = (int32_t)nd2;
                insert(nd2);
                ++since;
            }
        }
        if (on_len < n_nodes) {
            if (since >= block_dim || on_i == on_len - 1) {
                if (dead > n_rem) compact();
                int64_t last = ordered[on_len - 1];
                const double* p1 = xyz + 3 * last;
                double min_dist = 1.0e+100;
                int64_t min_nd = -1;
                int64_t len = n_rem + dead;
                for (int64_t r = 0; r < len; ++r) {
                    int32_t i = remaining[r];
                    if (i < 0) continue;
                    const double* p2 = xyz + 3 * (int64_t)i;
                    double dx = p1[0] - p2[0], dy = p1[1] - p2[1], dz = p1[2] - p2[2];
                    double d = std::sqrt(dx * dx + dy * dy + dz * dz);
                    if (d < min_dist) {
                        min_dist = d;
                        min_nd = i;
                    }
                }
                if (min_nd < 0) return ASEND_ERR_NUMERIC;  // non-finite coordinates
                ordered[on_len++] = (int32_t)min_nd;
                nd1 = min_nd;
                insert(min_nd);
                ++since;
                on_i = on_len - 1;
                if (since >= block_dim) since = 0;
            } else {
                ++on_i;
                nd1 = ordered[on_i];
            }
        }
    }
    for (int64_t i = 0; i < 6 * n_nodes; ++i) dof_index[i] = -1;
    int64_t d = 0;
    for (int64_t k = 0; k < n_nodes; ++k) {
        int64_t n = ordered[k];
        for (int c = 0; c < ndof[n]; ++c) dof_index[6 * n + c] = d++;
    }
    if (n_dof_out) *n_dof_out = d;
    return ASEND_OK;
}

// ---------------------------------------------------------------------------
// Hex shape functions and an FP64 Jacobian for the traction distribution.
static const double kHexNd[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                    {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
static const double kGp = 0.5773502691896257;

static double hex_detj(const double* X /*[8][3]*/, const double* s, double* N) {
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < 8; ++a) {
        double fx = 1.0 + kHexNd[a][0] * s[0], fy = 1.0 + kHexNd[a][1] * s[1], fz = 1.0 + kHexNd[a][2] * s[2];
        N[a] = 0.125 * fx * fy * fz;
        double d[3] = {0.125 * kHexNd[a][0] * fy * fz, 0.125 * fx * kHexNd[a][1] * fz,
                       0.125 * fx * fy * kHexNd[a][2]};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) J[i][j] += X[3 * a + i] * d[j];
    }
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

// hex face local nodes (S/element_meth.rs:537-590)
static const int kHexFace[6][4] = {{3, 2, 1, 0}, {4, 5, 6, 7}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};

// on_surf[6 e + k]: face k of hex element e lies on the surface, i.e. no
// other face has the same node set (S/model_prep.rs:510-554).  Faces are
// bucketed by their sorted nodes.
static std::vector<uint8_t> hex_surface_faces(int64_t n_el, const int32_t* conn) {
    const auto& fl = kHexFace;
    struct Key {
        int32_t n[4];
        bool operator<(const Key& o) const { return std::lexicographical_compare(n, n + 4, o.n, o.n + 4); }
        bool operator==(const Key& o) const { return std::equal(n, n + 4, o.n); }
    };
    std::vector<std::pair<Key, int64_t>> keys;
    keys.reserve(6 * n_el);
    for (int64_t e = 0; e < n_el; ++e)
        for (int k = 0; k < 6; ++k) {
            Key key;
            for (int i = 0; i < 4; ++i) key.n[i] = conn[8 * e + fl[k][i]];
            std::sort(key.n, key.n + 4);
            keys.push_back({key, 6 * e + k});
        }
    std::sort(keys.begin(), keys.end(), [](const std::pair<Key, int64_t>& a, const std::pair<Key, int64_t>& b) {
        return a.first < b.first;
    });
    std::vector<uint8_t> on_surf(6 * n_el, 1);
    for (size_t i = 0; i + 1 < keys.size(); ++i)
        if (keys[i].first == keys[i + 1].first) {
            on_surf[keys[i].second] = 0;
            on_surf[keys[i + 1].second] = 0;
        }
    return on_surf;
}

// surfaceTraction (traction = load) and surfacePressure (traction = -p n of
// each face, pressure != nullptr) share the face selection and distribution
static int surface_load(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                        const int64_t* dof_index, int64_t n_set, const int32_t* el_set, const double* load,
                        const double* pressure, const double* norm_dir_in, double norm_tol_deg, int64_t n_dof,
                        double* f) {
    if (elem_type != ASEND_BRICK8 && elem_type != ASEND_BRICKIM) return ASEND_ERR_UNSUPPORTED;
    // the load's normal direction is stored normalised (Load::set_norm_dir, S/load_meth.rs:22-32)
    const double nd2 = norm_dir_in[0] * norm_dir_in[0] + norm_dir_in[1] * norm_dir_in[1] + norm_dir_in[2] * norm_dir_in[2];
    if (!(nd2 > 0.0)) return ASEND_ERR_ARG;
    const double ndm = 1.0 / std::sqrt(nd2);
    const double norm_dir[3] = {ndm * norm_dir_in[0], ndm * norm_dir_in[1], ndm * norm_dir_in[2]};
    const auto& fl = kHexFace;
    const std::vector<uint8_t> on_surf = hex_surface_faces(n_el, conn);
    const double cos_tol = std::cos(0.017453292519943295 * norm_tol_deg);
    for (int64_t si = 0; si < n_set; ++si) {
        int64_t e = el_set[si];
        if (e < 0 || e >= n_el) return ASEND_ERR_ARG;
        const int32_t* c = conn + 8 * e;
        double X[24];
        for (int a = 0; a < 8; ++a)
            for (int k = 0; k < 3; ++k) X[3 * a + k] = xyz[3 * (int64_t)c[a] + k];
        double acc[8][3];
        double ld[8][3];
        double tot[3] = {0, 0, 0};
        bool any = false;
        std::memset(acc, 0, sizeof(acc));
        std::memset(ld, 0, sizeof(ld));
        for (int k = 0; k < 6; ++k) {
            if (!on_surf[6 * e + k]) continue;
            const int* l = fl[k];
            double v1[3], v2[3], nrm[3];
            for (int i = 0; i < 3; ++i) {
                v1[i] = X[3 * l[2] + i] - X[3 * l[0] + i];
                v2[i] = X[3 * l[3] + i] - X[3 * l[1] + i];
            }
            nrm[0] = v1[1] * v2[2] - v1[2] * v2[1];
            nrm[1] = v1[2] * v2[0] - v1[0] * v2[2];
            nrm[2] = v1[0] * v2[1] - v1[1] * v2[0];
            double a2 = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            double dp = (norm_dir[0] * nrm[0] + norm_dir[1] * nrm[1] + norm_dir[2] * nrm[2]) / a2;
            if (!(dp > cos_tol)) continue;
            any = true;
            double trac[3];
            for (int d = 0; d < 3; ++d) trac[d] = pressure ? -(*pressure) * (nrm[d] / a2) : load[d];
            const double inp = std::sqrt(trac[0] * trac[0] + trac[1] * trac[1] + trac[2] * trac[2]);
            for (int i = 0; i < 4; ++i)
                for (int d = 0; d < 3; ++d) acc[l[i]][d] = trac[d];
            // traction spread like a unit-density mass load over the element,
            // kept on the face nodes, then rescaled to area*|t|
            // (S/element_equations.rs:1970-2043)
            std::memset(ld, 0, sizeof(ld));
            for (int gp = 0; gp < 8; ++gp) {
                double s[3] = {(gp & 1) ? kGp : -kGp, (gp & 2) ? kGp : -kGp, (gp & 4) ? kGp : -kGp};
                double N[8];
                double w = hex_detj(X, s, N);
                double aip[3] = {0, 0, 0};
                for (int b = 0; b < 8; ++b)
                    for (int d = 0; d < 3; ++d) aip[d] += acc[b][d] * N[b];
                for (int a = 0; a < 8; ++a)
                    for (int d = 0; d < 3; ++d) ld[a][d] += N[a] * aip[d] * w;
            }
            for (int a = 0; a < 8; ++a) {
                bool in = (a == l[0] || a == l[1] || a == l[2] || a == l[3]);
                for (int d = 0; d < 3; ++d) {
                    if (!in) ld[a][d] = 0.0;
                    tot[d] += ld[a][d];
                }
            }
            double vm = std::sqrt(tot[0] * tot[0] + tot[1] * tot[1] + tot[2] * tot[2]);
            double sc = 0.5 * a2 * inp / vm;
            for (int a = 0; a < 8; ++a)
                for (int d = 0; d < 3; ++d) ld[a][d] *= sc;
        }
        if (!any) continue;  // element without a loaded face contributes nothing
        for (int a = 0; a < 8; ++a)
            for (int d = 0; d < 3; ++d) {
                int64_t g = dof_index[6 * (int64_t)c[a] + d];
                if (g < 0 || g >= n_dof) return ASEND_ERR_ARG;
                f[g] += ld[a][d];
            }
    }
    (void)n_nodes;
    return ASEND_OK;
}

int surface_traction(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                     const int64_t* dof_index, int64_t n_set, const int32_t* el_set, const double* load,
                     const double* norm_dir, double norm_tol_deg, int64_t n_dof, double* f) {
    return surface_load(elem_type, n_nodes, xyz, n_el, conn, dof_index, n_set, el_set, load, nullptr, norm_dir,
                        norm_tol_deg, n_dof, f);
}

int surface_pressure(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                     const int64_t* dof_index, int64_t n_set, const int32_t* el_set, double pressure,
                     const double* norm_dir, double norm_tol_deg, int64_t n_dof, double* f) {
    return surface_load(elem_type, n_nodes, xyz, n_el, conn, dof_index, n_set, el_set, nullptr, &pressure, norm_dir,
                        norm_tol_deg, n_dof, f);
}

// surfaceFlux element loads of hex elements (Element::get_app_therm_load_dfd0,
// S/element_equations.rs:2113-2150): the face test adds norm_dir . n of every
// surface face of the element to one running sum (never reset between
// faces); a face passes while that sum exceeds cos(norm_tol).  A passing
// face sets the nodal field to q on its nodes (kept across the element's
// faces), takes the unit-capacity heat-rate integral
// sum_gp w detJ N_a sum_b N_b T_b (get_rtm_dfd0 without the material's
// specific heat), keeps the face's nodes and rescales them to q times the
// face area.  The element adds the vector of its last passing face.
int surface_flux(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                 const int64_t* dof_index, int64_t n_set, const int32_t* el_set, double q, const double* norm_dir_in,
                 double norm_tol_deg, int64_t n_dof, double* f) {
    if (elem_type != ASEND_BRICK8 && elem_type != ASEND_BRICKIM) return ASEND_ERR_UNSUPPORTED;
    const double nd2 = norm_dir_in[0] * norm_dir_in[0] + norm_dir_in[1] * norm_dir_in[1] + norm_dir_in[2] * norm_dir_in[2];
    if (!(nd2 > 0.0)) return ASEND_ERR_ARG;
    const double ndm = 1.0 / std::sqrt(nd2);
    const double norm_dir[3] = {ndm * norm_dir_in[0], ndm * norm_dir_in[1], ndm * norm_dir_in[2]};
    const auto& fl = kHexFace;
    const std::vector<uint8_t> on_surf = hex_surface_faces(n_el, conn);
    const double cos_tol = std::cos(0.017453292519943295 * norm_tol_deg);
    for (int64_t si = 0; si < n_set; ++si) {
        const int64_t e = el_set[si];
        if (e < 0 || e >= n_el) return ASEND_ERR_ARG;
        const int32_t* c = conn + 8 * e;
        double X[24];
        for (int a = 0; a < 8; ++a)
            for (int k = 0; k < 3; ++k) X[3 * a + k] = xyz[3 * (int64_t)c[a] + k];
        double T[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        double ld[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        double dp = 0.0;
        bool any = false;
        for (int k = 0; k < 6; ++k) {
            if (!on_surf[6 * e + k]) continue;
            const int* l = fl[k];
            double v1[3], v2[3], nrm[3];
            for (int i = 0; i < 3; ++i) {
                v1[i] = X[3 * l[2] + i] - X[3 * l[0] + i];
                v2[i] = X[3 * l[3] + i] - X[3 * l[1] + i];
            }
            nrm[0] = v1[1] * v2[2] - v1[2] * v2[1];
            nrm[1] = v1[2] * v2[0] - v1[0] * v2[2];
            nrm[2] = v1[0] * v2[1] - v1[1] * v2[0];
            const double a2 = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            for (int i = 0; i < 3; ++i) dp += norm_dir[i] * (nrm[i] / a2);
            if (!(dp > cos_tol)) continue;
            any = true;
            for (int i = 0; i < 4; ++i) T[l[i]] = q;
            for (int a = 0; a < 8; ++a) ld[a] = 0.0;
            for (int gp = 0; gp < 8; ++gp) {
                const double s[3] = {(gp & 1) ? kGp : -kGp, (gp & 2) ? kGp : -kGp, (gp & 4) ? kGp : -kGp};
                double N[8];
                const double w = hex_detj(X, s, N);
                double tp = 0.0;
                for (int b = 0; b < 8; ++b) tp += N[b] * T[b];
                for (int a = 0; a < 8; ++a) ld[a] += N[a] * tp * w;
            }
            double tot = 0.0;
            for (int a = 0; a < 8; ++a) {
                if (a != l[0] && a != l[1] && a != l[2] && a != l[3]) ld[a] = 0.0;
                tot += ld[a];
            }
            const double sc = q * (0.5 * a2) / tot;
            for (int a = 0; a < 8; ++a) ld[a] *= sc;
        }
        if (!any) continue;  // element without a passing face contributes nothing
        for (int a = 0; a < 8; ++a) {
            const int64_t g = dof_index[6 * (int64_t)c[a]];
            if (g < 0 || g >= n_dof) return ASEND_ERR_ARG;
            f[g] += ld[a];
        }
    }
    (void)n_nodes;
    return ASEND_OK;
}

// Thermal strain load of hex elements (the expansion term of
// get_ruk_dfd0, S/element_equations.rs:272-298: the strain at each Gauss
// point loses alpha times the interpolated nodal temperature, no reference
// temperature): f_a = sum_gp w detJ S g_a with S the stress tensor of
// C (alpha T_gp), C isotropic from (E_e, nu_e) or, when cmat is given, the
// element's 6 x 6 matrix times E_e; alpha: 6 per element (strain order 11,
// 22, 33, 12, 13, 23, engineering shear); temp: per node label.
int thermal_strain_load(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                        const int64_t* dof_index, const double* E, const double* nu, const double* cmat,
                        const int32_t* cmat_idx, const double* alpha, const double* temp, int64_t n_dof, double* f) {
    if (elem_type != ASEND_BRICK8) return ASEND_ERR_UNSUPPORTED;
    for (int64_t e = 0; e < n_el; ++e) {
        const int32_t* c = conn + 8 * e;
        const double* al = alpha + 6 * e;
        if (al[0] == 0.0 && al[1] == 0.0 && al[2] == 0.0 && al[3] == 0.0 && al[4] == 0.0 && al[5] == 0.0) continue;
        double X[24], T[8];
        for (int a = 0; a < 8; ++a) {
            if (c[a] < 0 || c[a] >= n_nodes) return ASEND_ERR_ARG;
            T[a] = temp[c[a]];
            for (int k = 0; k < 3; ++k) X[3 * a + k] = xyz[3 * (int64_t)c[a] + k];
        }
        double Cm[36];
        if (cmat) {
            const double* src = cmat + 36 * (cmat_idx ? (int64_t)cmat_idx[e] : 0);
            for (int i = 0; i < 36; ++i) Cm[i] = E[e] * src[i];
        } else {
            const double lam = E[e] * nu[e] / ((1.0 + nu[e]) * (1.0 - 2.0 * nu[e])), mu = E[e] / (2.0 * (1.0 + nu[e]));
            for (int i = 0; i < 36; ++i) Cm[i] = 0.0;
            for (int i = 0; i < 3; ++i) {
                for (int j = 0; j < 3; ++j) Cm[6 * i + j] = lam;
                Cm[7 * i] += 2.0 * mu;
                Cm[7 * (i + 3)] = mu;
            }
        }
        double ld[8][3];
        std::memset(ld, 0, sizeof(ld));
        for (int gp = 0; gp < 8; ++gp) {
            const double s[3] = {(gp & 1) ? kGp : -kGp, (gp & 2) ? kGp : -kGp, (gp & 4) ? kGp : -kGp};
            double N[8], J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, dn[8][3];
            for (int a = 0; a < 8; ++a) {
                const double fx = 1.0 + kHexNd[a][0] * s[0], fy = 1.0 + kHexNd[a][1] * s[1],
                             fz = 1.0 + kHexNd[a][2] * s[2];
                N[a] = 0.125 * fx * fy * fz;
                dn[a][0] = 0.125 * kHexNd[a][0] * fy * fz;
                dn[a][1] = 0.125 * fx * kHexNd[a][1] * fz;
                dn[a][2] = 0.125 * fx * fy * kHexNd[a][2];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) J[i][j] += X[3 * a + i] * dn[a][j];
            }
            const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2],
                         c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
            const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
            const double Ji[3][3] = {{c00 / det, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det,
                                      (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det},
                                     {c01 / det, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det,
                                      (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det},
                                     {c02 / det, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det,
                                      (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det}};
            double Tg = 0.0;
            for (int a = 0; a < 8; ++a) Tg += N[a] * T[a];
            double sv[6];
            for (int i = 0; i < 6; ++i) {
                sv[i] = 0.0;
                for (int j = 0; j < 6; ++j) sv[i] += Cm[6 * i + j] * al[j] * Tg;
            }
            const double S[3][3] = {{sv[0], sv[3], sv[4]}, {sv[3], sv[1], sv[5]}, {sv[4], sv[5], sv[2]}};
            for (int a = 0; a < 8; ++a) {
                double g[3];
                for (int j = 0; j < 3; ++j) g[j] = dn[a][0] * Ji[0][j] + dn[a][1] * Ji[1][j] + dn[a][2] * Ji[2][j];
                for (int i = 0; i < 3; ++i) ld[a][i] += det * (S[i][0] * g[0] + S[i][1] * g[1] + S[i][2] * g[2]);
            }
        }
        for (int a = 0; a < 8; ++a)
            for (int k = 0; k < 3; ++k) {
                const int64_t g = dof_index[6 * (int64_t)c[a] + k];
                if (g < 0 || g >= n_dof) return ASEND_ERR_ARG;
                f[g] += ld[a][k];
            }
    }
    return ASEND_OK;
}

// bodyForce / gravitational / centrifugal element loads of hex elements
// (Element::get_app_load_dfd0, S/element_equations.rs:1876-1969): a nodal
// acceleration field pushed through the consistent mass integral
// f_a = sum_gp detJ N_a rho sum_b N_b acc_b (get_rum_dfd0) over the element's
// 8 Gauss points.  bodyForce: acc = b at every node, rho = 1, then the
// element's vector is rescaled to |b| V / |sum_a f_a|; gravitational:
// acc = g, rho = the element density; centrifugal: acc = omega^2 times the
// component of (centroid - center) normal to the unit axis, rho = the
// element density.
int body_load(int elem_type, int kind, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
              const int64_t* dof_index, int64_t n_set, const int32_t* el_set, const double* rho, const double* load,
              const double* center, const double* axis_in, double ang_vel, int64_t n_dof, double* f) {
    if (elem_type != ASEND_BRICK8 && elem_type != ASEND_BRICKIM) return ASEND_ERR_UNSUPPORTED;
    if (kind < 0 || kind > 2) return ASEND_ERR_ARG;
    if (kind > 0 && !rho) return ASEND_ERR_ARG;
    double axis[3] = {0.0, 0.0, 0.0};
    if (kind == 2) {
        // Load::set_axis normalises (S/load_meth.rs:41-51)
        const double a2 = axis_in[0] * axis_in[0] + axis_in[1] * axis_in[1] + axis_in[2] * axis_in[2];
        if (!(a2 > 0.0)) return ASEND_ERR_ARG;
        for (int k = 0; k < 3; ++k) axis[k] = axis_in[k] / std::sqrt(a2);
    }
    for (int64_t si = 0; si < n_set; ++si) {
        const int64_t e = el_set[si];
        if (e < 0 || e >= n_el) return ASEND_ERR_ARG;
        const int32_t* c = conn + 8 * e;
        double X[24];
        for (int a = 0; a < 8; ++a)
            for (int k = 0; k < 3; ++k) X[3 * a + k] = xyz[3 * (int64_t)c[a] + k];
        double acc[3];
        double dens = 1.0;
        if (kind == 0) {
            for (int k = 0; k < 3; ++k) acc[k] = load[k];
        } else if (kind == 1) {
            for (int k = 0; k < 3; ++k) acc[k] = load[k];
            dens = rho[e];
        } else {
            double cen[3] = {0, 0, 0}, dp = 0.0;
            for (int k = 0; k < 3; ++k) {
                for (int a = 0; a < 8; ++a) cen[k] += X[3 * a + k];
                cen[k] = cen[k] * 0.125 - center[k];
                dp += axis[k] * cen[k];
            }
            for (int k = 0; k < 3; ++k) acc[k] = ang_vel * ang_vel * (cen[k] - axis[k] * dp);
            dens = rho[e];
        }
        // every node carries the same acceleration, so sum_b N_b acc_b = acc
        double ld[8][3];
        std::memset(ld, 0, sizeof(ld));
        double vol = 0.0;
        for (int gp = 0; gp < 8; ++gp) {
            const double s[3] = {(gp & 1) ? kGp : -kGp, (gp & 2) ? kGp : -kGp, (gp & 4) ? kGp : -kGp};
            double N[8];
            const double w = hex_detj(X, s, N);
            double sn = 0.0;
            for (int b = 0; b < 8; ++b) sn += N[b];
            vol += w;
            for (int a = 0; a < 8; ++a)
                for (int k = 0; k < 3; ++k) ld[a][k] += N[a] * dens * (sn * acc[k]) * w;
        }
        if (kind == 0) {
            double tot[3] = {0, 0, 0};
            for (int a = 0; a < 8; ++a)
                for (int k = 0; k < 3; ++k) tot[k] += ld[a][k];
            const double inp = std::sqrt(acc[0] * acc[0] + acc[1] * acc[1] + acc[2] * acc[2]);
            const double vm = std::sqrt(tot[0] * tot[0] + tot[1] * tot[1] + tot[2] * tot[2]);
            const double sc = (vm > 0.0) ? inp * vol / vm : 0.0;
            for (int a = 0; a < 8; ++a)
                for (int k = 0; k < 3; ++k) ld[a][k] *= sc;
        }
        for (int a = 0; a < 8; ++a)
            for (int k = 0; k < 3; ++k) {
                const int64_t g = dof_index[6 * (int64_t)c[a] + k];
                if (g < 0 || g >= n_dof) return ASEND_ERR_ARG;
                f[g] += ld[a][k];
            }
    }
    (void)n_nodes;
    return ASEND_OK;
}

// splitmix64 (Steele, Lea, Flood 2014) + Box-Muller, cosine branch.
static inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int lognormal_moduli(uint64_t seed, int64_t n, double median, double sigma_ln, double* out) {
    if (n < 0 || !out) return ASEND_ERR_ARG;
    uint64_t s = seed;
    const double two_pi = 6.283185307179586;
    for (int64_t i = 0; i < n; ++i) {
        // u1 in (0,1], u2 in [0,1): 53-bit uniforms
        double u1 = ((splitmix64(s) >> 11) + 1) * (1.0 / 9007199254740992.0);
        double u2 = (splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
        double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(two_pi * u2);
        out[i] = median * std::exp(sigma_ln * z);
    }
    return ASEND_OK;
}

// The reference builds an isotropic stiffness by inverting the compliance
// S(E, nu, G = E/(2(1+nu))): the 6x6 of get_solid_stiff_dfd0
// (S/element_properties.rs:728-759) for solids, the 3x3 plane-stress layer
// matrix of get_layer_q_dfd0 (:196-210) for shells.  The inverse is the
// Givens QR of q_rfactor (S/matrix_functions.rs:167-215): before each
// rotation of column i1 against a lower row, a diagonal entry of magnitude
// below 1e-12 (S/constants.rs:13) is replaced by 1e-12; the last column has no
// rotation and is never checked.  Once that fires, the result is no longer
// S^-1, while the device integrates the exact isotropic stiffness.  The
// checks below follow the rotations in the reference's order and arithmetic
// (theta = atan(b/a), cos and sin of theta) so that the decision at the
// threshold is the reference's.
static void iso_compliance(int n, double E, double nu, double* m) {
    const double G = E / (2.0 * (1.0 + nu));
    for (int i = 0; i < n * n; ++i) m[i] = 0.0;
    if (n == 6) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) m[6 * i + j] = (i == j) ? 1.0 / E : -nu / E;
        m[21] = m[28] = m[35] = 1.0 / G;
    } else {
        m[0] = m[4] = 1.0 / E;
        m[1] = m[3] = -nu / E;
        m[8] = 1.0 / G;
    }
}

// true when some checked pivot of the QR of m (n x n) is below 1e-12
static bool qr_pivot_clamps(int n, double* m) {
    for (int i1 = 0; i1 < n; ++i1)
        for (int i2 = i1 + 1; i2 < n; ++i2) {
            const double a = m[n * i1 + i1];
            if (std::fabs(a) < 1e-12) return true;
            const double th = std::atan(m[n * i2 + i1] / a), c = std::cos(th), s = std::sin(th);
            for (int i3 = i1; i3 < n; ++i3) {
                const double p = m[n * i1 + i3], q = m[n * i2 + i3];
                m[n * i1 + i3] = c * p + s * q;
                m[n * i2 + i3] = -s * p + c * q;
            }
        }
    return false;
}

// Smallest checked pivot magnitude of the QR of S(1, nu): every pivot of
// S(E, nu) is this scaled by 1/E up to rounding, which screens out the
// elements far from the threshold without trigonometry.
static double min_scaled_pivot(int n, double nu) {
    double m[36];
    iso_compliance(n, 1.0, nu, m);
    double lo = 1e300;
    for (int i1 = 0; i1 < n; ++i1)
        for (int i2 = i1 + 1; i2 < n; ++i2) {
            const double a = m[n * i1 + i1], b = m[n * i2 + i1];
            lo = std::min(lo, std::fabs(a));
            const double r = std::hypot(a, b);
            if (r == 0.0) continue;
            const double c = a / r, s = b / r;
            for (int i3 = i1; i3 < n; ++i3) {
                const double p = m[n * i1 + i3], q = m[n * i2 + i3];
                m[n * i1 + i3] = c * p + s * q;
                m[n * i2 + i3] = -s * p + c * q;
            }
        }
    return lo;
}

int64_t compliance_clamp_count(int elem_type, int64_t n_el, const double* E, const double* nu, int64_t* first) {
    const int n = (elem_type == ASEND_SHELL4) ? 3 : 6;
    int64_t count = 0;
    if (first) *first = -1;
    double last_nu = std::nan(""), lo = 0.0;
    for (int64_t e = 0; e < n_el; ++e) {
        if (!(nu[e] == last_nu)) {
            last_nu = nu[e];
            lo = min_scaled_pivot(n, nu[e]);
        }
        // pivots are lo/E to within a few ulps: only elements within a
        // factor 2 of the threshold take the exact replay
        if (!(std::fabs(E[e]) * 2e-12 < lo)) {
            double m[36];
            iso_compliance(n, E[e], nu[e], m);
            if (qr_pivot_clamps(n, m)) {
                if (count == 0 && first) *first = e;
                ++count;
            }
        }
    }
    return count;
}

}  // namespace asend

extern "C" {

int asend_reorder_nodes(int64_t n_nodes, const double* xyz, int64_t n_el, int nn, int dof_per_nd,
                        const int32_t* conn, int64_t n_const, const int64_t* cterm_ptr, const int64_t* term_set_ptr,
                        const int32_t* set_nodes, int64_t block_dim, int32_t* ordered, int64_t* dof_index,
                        int64_t* n_dof_out) {
    return asend::reorder_nodes(n_nodes, xyz, n_el, nn, dof_per_nd, conn, n_const, cterm_ptr, term_set_ptr,
                                set_nodes, block_dim, ordered, dof_index, n_dof_out);
}

int64_t asend_block_dim_auto(int64_t n_nodes) {
    int64_t b = 6;
    while (b * b < n_nodes) b += 6;
    return b;
}

int asend_surface_traction(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                           const int64_t* dof_index, int64_t n_set, const int32_t* el_set, const double* load,
                           const double* norm_dir, double norm_tol_deg, int64_t n_dof, double* f) {
    return asend::surface_traction(elem_type, n_nodes, xyz, n_el, conn, dof_index, n_set, el_set, load, norm_dir,
                                   norm_tol_deg, n_dof, f);
}

int asend_lognormal_moduli(uint64_t seed, int64_t n, double median, double sigma_ln, double* out) {
    return asend::lognormal_moduli(seed, n, median, sigma_ln, out);
}

int asend_surface_pressure(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                           const int64_t* dof_index, int64_t n_set, const int32_t* el_set, double pressure,
                           const double* norm_dir, double norm_tol_deg, int64_t n_dof, double* f) {
    return asend::surface_pressure(elem_type, n_nodes, xyz, n_el, conn, dof_index, n_set, el_set, pressure, norm_dir,
                                   norm_tol_deg, n_dof, f);
}

int asend_surface_flux(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                       const int64_t* dof_index, int64_t n_set, const int32_t* el_set, double q,
                       const double* norm_dir, double norm_tol_deg, int64_t n_dof, double* f) {
    if (!xyz || !conn || !dof_index || (n_set > 0 && !el_set) || !norm_dir || !f) return ASEND_ERR_ARG;
    return asend::surface_flux(elem_type, n_nodes, xyz, n_el, conn, dof_index, n_set, el_set, q, norm_dir,
                               norm_tol_deg, n_dof, f);
}

int asend_thermal_strain_load(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                              const int64_t* dof_index, const double* E, const double* nu, const double* cmat,
                              const int32_t* cmat_idx, const double* alpha, const double* temp, int64_t n_dof,
                              double* f) {
    if (!xyz || !conn || !dof_index || !E || (!nu && !cmat) || !alpha || !temp || !f) return ASEND_ERR_ARG;
    return asend::thermal_strain_load(elem_type, n_nodes, xyz, n_el, conn, dof_index, E, nu, cmat, cmat_idx, alpha,
                                      temp, n_dof, f);
}

int asend_body_load(int elem_type, int kind, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                    const int64_t* dof_index, int64_t n_set, const int32_t* el_set, const double* rho,
                    const double* load, const double* center, const double* axis, double angular_velocity,
                    int64_t n_dof, double* f) {
    return asend::body_load(elem_type, kind, n_nodes, xyz, n_el, conn, dof_index, n_set, el_set, rho, load, center,
                            axis, angular_velocity, n_dof, f);
}

int asend_compliance_clamp_check(int elem_type, int64_t n_el, const double* E, const double* nu,
                                 int64_t* n_clamped, int64_t* first) {
    if (elem_type != ASEND_BRICK8 && elem_type != ASEND_BRICKIM && elem_type != ASEND_SHELL4) return ASEND_ERR_ARG;
    if (n_el < 0 || (n_el > 0 && (!E || !nu)) || !n_clamped) return ASEND_ERR_ARG;
    *n_clamped = asend::compliance_clamp_count(elem_type, n_el, E, nu, first);
    return ASEND_OK;
}

const char* asend_version(void) { return "asend-mi355x 0.1 (gfx950)"; }

}  // extern "C"
