// asend_run_job <job.yaml>: the reference's solver executable
// (S/main.rs:55-70) for linear elastic jobs, running the solve path through
// the C ABI of libasend_gpu.so on the GPU.
//
// Job, model, constraint, load, design-variable and objective files are read
// with the reference's line rules (read_input_line, S/model_input.rs:34-119)
// and keys (read_job :121-310, read_model_input :312-860,
// read_constraint_input :872-969, read_load_input :981-1123,
// read_des_var_input :1218-1296, read_objective_input :1298-1375,
// read_des_var_values :1377-1401); commands run in file order as in
// execute_job (S/model_meth.rs:8-138).  Supported subset:
//   - one element type per model: brick8, brickIM or shell4 (one layer);
//     isotropic materials (E, nu, G = E/(2(1+nu)) after design variables);
//   - displacement constraints with one term of coefficient +-1 and rhs 0
//     (clamps, applied by the penalty method);
//   - loads nodalForce, and for solid elements surfaceTraction,
//     surfacePressure, bodyForce, gravitational and centrifugal, each with
//     its activeTime window;
//   - solve, linear elastic: static (one solve per staticLoadTime) with
//     solverMethod direct (default: the device LDL^T factor of the whole
//     operator) or iterative (conj_grad_sparse with the device block LDL^T
//     preconditioner and the reference stop rule, convergenceTol,
//     maxIterations, 0 meaning the DOF count as in the reference); or
//     dynamic (Newmark with newmarkBeta, newmarkGamma, timeStep, simPeriod,
//     rayDampCM, rayDampCK, material damping matrices on brick8; consistent
//     mass), each step solved on the device;
//     saveSolnHist writes the reference's solnTStep<i>.out state files into
//     solnHistDir;
//   - objective terms of category displacement, and on brick8 stress,
//     strain and strainEnergy at the element centres, with operators
//     powerNorm, volumeIntegral and volumeAverage (static solves);
//   - design variables of category modulus, poissonRatio, shearModulus,
//     density and nodeCoord (brick8; nodeCoord with nodal forces only) for
//     calcObjGradient;
//   - writeNodeResults (displacement, velocity, acceleration, reactionForce;
//     timeSteps read back from the state files) and writeObjective (with
//     the _grad file), numbers formatted as the reference's {:.12e}.
// Anything else stops the job with an error message, as the reference's
// panics do.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "asend_gpu.h"

namespace {

[[noreturn]] void die(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::fprintf(stderr, "Error: ");
    std::vfprintf(stderr, fmt, ap);
    std::fprintf(stderr, "\n");
    va_end(ap);
    std::exit(1);
}

// ---------------------------------------------------------------- input lines
// Heading levels are assigned by the column of a key's first character: the
// first key sets level 0, a key at a new column the next free level, a key at
// a known column that level (clearing the deeper ones).  Values after the
// colon, or after "- " on lines without a key, are split on " ,[]\t".
struct LineReader {
    std::string head[4];
    size_t col[4] = {0, 0, 0, 0};

    static void tokens(const std::string& s, size_t from, std::vector<std::string>& out) {
        const char* sep = " ,[]\t\n\r";
        size_t i = s.find_first_not_of(sep, from);
        while (i != std::string::npos) {
            size_t j = s.find_first_of(sep, i);
            if (j == std::string::npos) j = s.size();
            out.push_back(s.substr(i, j - i));
            i = s.find_first_not_of(sep, j);
        }
    }

    // returns true when the line holds a key (the headings were updated)
    bool read(std::string line, std::vector<std::string>& data) {
        data.clear();
        size_t h = line.find('#');
        if (h != std::string::npos) line = line.substr(0, h);
        line += " ";
        size_t c = line.find(':');
        if (c != std::string::npos) {
            size_t i2 = line.find_first_not_of(" -\n\t");
            std::string word = (i2 < c) ? line.substr(i2, c - i2) : std::string();
            int lev;
            if (head[0].empty() || col[0] == i2) lev = 0;
            else if (head[1].empty() || col[1] == i2) lev = 1;
            else if (head[2].empty() || col[2] == i2) lev = 2;
            else lev = 3;
            head[lev] = word;
            col[lev] = i2;
            for (int k = lev + 1; k < 4; ++k) {
                head[k].clear();
                col[k] = 0;
            }
            tokens(line, c + 1, data);
            return true;
        }
        size_t d = line.find("- ");
        if (d != std::string::npos) tokens(line, d + 1, data);
        return false;
    }
};

std::vector<std::string> read_lines(const std::string& path, const char* what) {
    std::ifstream in(path);
    if (!in) die("could not open %s file: %s", what, path.c_str());
    std::vector<std::string> lines;
    std::string s;
    while (std::getline(in, s)) lines.push_back(s);
    return lines;
}

double to_d(const std::string& s) {
    char* end = nullptr;
    double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str()) die("could not read '%s' as a number", s.c_str());
    return v;
}
long to_i(const std::string& s) { return (long)std::llround(to_d(s)); }
bool yes(const std::string& s) { return s.find("yes") != std::string::npos; }

// Rust's {:.12e}: 12 decimals, exponent without sign padding or leading zeros
std::string fmt_e12(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.12e", v);
    std::string s(buf);
    size_t e = s.find('e');
    std::string mant = s.substr(0, e);
    int ex = std::atoi(s.c_str() + e + 1);
    return mant + "e" + std::to_string(ex);
}

// Rust's {} for f64: the shortest digits that read back to the same double,
// written out in plain decimal notation without an exponent
std::string fmt_display(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    if (v == 0.0) return std::signbit(v) ? "-0" : "0";
    char buf[64];
    for (int p = 1; p <= 17; ++p) {
        std::snprintf(buf, sizeof buf, "%.*e", p - 1, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    std::string s(buf);
    const bool neg = s[0] == '-';
    if (neg) s = s.substr(1);
    const size_t e = s.find('e');
    const int ex = std::atoi(s.c_str() + e + 1);
    std::string digits;
    for (char ch : s.substr(0, e))
        if (ch != '.') digits += ch;
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    const int point = ex + 1;  // digits before the decimal point
    std::string out;
    if (point <= 0) out = "0." + std::string(-point, '0') + digits;
    else if (point >= (int)digits.size()) out = digits + std::string(point - digits.size(), '0');
    else out = digits.substr(0, point) + "." + digits.substr(point);
    return (neg ? "-" : "") + out;
}

// ---------------------------------------------------------------- the model
struct Section {
    std::string type, mat, el_set;
    // local axes as rows (Section::set_orientation, S/section_meth.rs:21-50)
    double ori[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    bool oriented = false;
    std::vector<std::string> lay_mat;
    std::vector<double> lay_t;
    double z_offset = 0.0;
};
struct Material {
    std::string name;
    double density = 0.0;
    double E[3] = {0, 0, 0}, nu[3] = {0, 0, 0}, G[3] = {0, 0, 0};
    bool stiffness = false;
    double stiff[36] = {};  // stiffness entries (Material::set_stiffness, symmetric)
    double damping[36] = {};  // Material::set_damping (symmetric, S/section_meth.rs:73-77)
    bool damped = false;
    double cond[6] = {0, 0, 0, 0, 0, 0};  // thermal conductivity k11 k22 k33 k12 k13 k23
    double spec_heat = 0.0;
    bool expands = false;                 // a nonzero thermal expansion coefficient
    double expansion[6] = {0, 0, 0, 0, 0, 0};
};
struct Term {
    std::string set;
    int dof = 1;
    double coef = 1.0;
};
struct Constraint {
    std::string type;
    std::vector<Term> terms;
    double rhs = 0.0;
};
struct Load {
    std::string type, node_set, el_set;
    double load[6] = {0, 0, 0, 0, 0, 0};
    double norm_dir[3] = {0, 0, 0};
    double norm_tol = 0.0;
    double center[3] = {0, 0, 0}, axis[3] = {0, 0, 0};
    double ang_vel = 0.0;
    double active[2] = {0.0, 1.0e100};  // activeTime (Load::active_time, S/load.rs:25)
    std::vector<double> vec;            // assembled load vector (reference DOF numbering)
};
struct DesignVar {
    std::string category, el_set, nd_set;
    int component = 1;
    std::vector<double> coefs;
    double value = 0.0;
};
struct ObjTerm {
    std::string category, optr, el_set, nd_set;
    int component = 1;
    double coef = 1.0, expnt = 1.0;
    std::vector<double> tgt;
    double value = 0.0;
};
struct Command {
    std::string cmd, file, method, node_set = "all", el_set = "all", position = "centroid";
    // modalAnalysis (JobCommand defaults, S/job.rs:92-112)
    std::string modal_type = "buckling";
    long num_modes = 10;
    double tgt_eval = 0.0;
    bool write_modes = true;
    // setSolnToMode
    std::string soln_field = "displacement";
    long mode = 0;
    double max_amp = 1.0;
    bool elastic = true, dynamic = false, nonlinear = false, thermal = false, write_grad = true;
    long block_dim = -1, max_it = 0, load_ramp = 1;
    double conv_tol = 1.0e-12;
    // dynamic and history keys (JobCommand defaults, S/job.rs:55-110)
    double time_step = 1.0, beta = 0.25, gamma = 0.5, sim_period = 1.0, ray_cm = 0.0, ray_ck = 0.0;
    bool save_hist = true, lump_mass = false;
    std::vector<double> load_times;
    std::vector<std::string> fields, include;
    std::vector<long> time_steps;
};
struct NamedSet {
    std::string name;
    std::vector<int32_t> labels;
};

struct Job {
    std::vector<Command> cmds;
    std::vector<double> xyz;  // 3 per node label
    std::vector<int> el_type;
    std::vector<std::vector<int32_t>> conn;
    std::vector<NamedSet> nsets, esets;
    std::map<std::string, size_t> ns_map, es_map;
    std::vector<Section> sections;
    std::vector<Material> materials;
    std::vector<Constraint> constraints;   // displacement
    std::vector<Constraint> tconstraints;  // temperature
    std::vector<Load> loads;               // elastic
    std::vector<Load> tloads;              // thermal
    std::vector<DesignVar> dvars;
    std::vector<ObjTerm> obj;

    // analysis state
    asend_ctx* ctx = nullptr;
    int etype = 0, dpn = 3;
    int64_t n_dof = 0;
    std::vector<int64_t> dof_index;  // 6 per node
    std::vector<double> E, nu, rho, thk;
    std::vector<int> mat_of_el;   // each element's material
    std::vector<int> sec_of_el;   // and section
    std::vector<double> cmats;    // global 6x6 stiffness per section when any is anisotropic (else empty)
    std::vector<int64_t> pen;
    std::vector<double> u;    // solution, reference DOF numbering
    std::vector<double> v, a; // velocity and acceleration (dynamic solves; zero for static)
    bool dyn_solved = false;  // the last solve was dynamic
    bool nl_solved = false;   // the last solve had nonlinearGeom
    std::vector<double> f;    // applied load
    std::vector<double> dLdD;
    bool solved = false, have_grad = false;
    int64_t solver_block_dim = 0;        // nodes per reordering block (2e9: direct)
    bool ldl_ok = true, bw_printed = false;  // LDL^T factor usable / maxBandwidth line printed
    Command solve_cmd;
    // heat conduction (thermal solves): its own context, the conductivities
    // of the elements, the node temperatures, and the penalised temperature
    // DOFs with coefficient x rhs of their constraints
    asend_ctx* tctx = nullptr;
    asend_ctx* cctx = nullptr;  // heat capacity alone (transient solves)
    std::vector<double> kc, temp, tdot, t3dot;
    std::vector<int64_t> tpen, tpen3;  // temperature DOFs; with their two companions
    std::vector<double> tpen_rhs;
    std::vector<double> t3;  // temperatures in the conduction context's DOFs
    bool tldl_ok = true, tbw_printed = false, thermal_solved = false;

    ~Job() {
        if (ctx) asend_ctx_destroy(ctx);
        if (tctx) asend_ctx_destroy(tctx);
        if (cctx) asend_ctx_destroy(cctx);
    }

    const std::vector<int32_t>& node_set(const std::string& n, bool fallback_all) {
        auto it = ns_map.find(n);
        if (it == ns_map.end()) {
            if (!fallback_all) die("there is no Node Set named %s", n.c_str());
            std::printf("Warning: there is no Node Set named %s. Defaulting to all nodes in writeNodeResults\n",
                        n.c_str());
            it = ns_map.find("all");
        }
        return nsets[it->second].labels;
    }
    const std::vector<int32_t>& element_set(const std::string& n) {
        auto it = es_map.find(n);
        if (it == es_map.end()) die("there is no Element Set named %s", n.c_str());
        return esets[it->second].labels;
    }

    void check(int rc, const char* what) {
        if (rc != ASEND_OK) die("%s failed (%d): %s", what, rc, ctx ? asend_last_error(ctx) : "");
    }
    void check_t(int rc, const char* what) {
        if (rc != ASEND_OK) die("%s failed (%d): %s", what, rc, tctx ? asend_last_error(tctx) : "");
    }

    // ------------------------------------------------------------ readers
    void read_job(const std::string& path) {
        LineReader lr;
        std::vector<std::string> d;
        for (const auto& line : read_lines(path, "job")) {
            lr.read(line, d);
            const std::string& k = lr.head[1];
            if (k == "command" && d.size() == 1) {
                cmds.push_back(Command());
                cmds.back().cmd = d[0];
                continue;
            }
            if (cmds.empty()) continue;
            Command& c = cmds.back();
            if (k == "fileName" && d.size() == 1) c.file = d[0];
            else if (k == "dynamic" && d.size() == 1) c.dynamic = yes(d[0]);
            else if (k == "elastic" && d.size() == 1) c.elastic = yes(d[0]);
            else if (k == "nonlinearGeom" && d.size() == 1) c.nonlinear = yes(d[0]);
            else if ((k == "thermal" || k == "diffusion") && d.size() == 1) c.thermal = yes(d[0]);
            else if (k == "solverBlockDim" && d.size() == 1) c.block_dim = to_i(d[0]);
            else if (k == "loadRampSteps" && d.size() == 1) c.load_ramp = to_i(d[0]);
            else if (k == "solverMethod" && d.size() == 1) c.method = d[0];
            else if (k == "maxIterations" && d.size() == 1) c.max_it = to_i(d[0]);
            else if (k == "convergenceTol" && d.size() == 1) c.conv_tol = to_d(d[0]);
            else if (k == "nodeSet" && d.size() == 1) c.node_set = d[0];
            else if (k == "elementSet" && d.size() == 1) c.el_set = d[0];
            else if (k == "type" && d.size() == 1) c.modal_type = d[0];
            else if (k == "numModes" && d.size() == 1) c.num_modes = to_i(d[0]);
            else if (k == "targetEigenvalue" && d.size() == 1) c.tgt_eval = to_d(d[0]);
            else if (k == "writeModes" && d.size() == 1) c.write_modes = (d[0] == "yes");
            else if (k == "solnField" && d.size() == 1) c.soln_field = d[0];
            else if (k == "mode" && d.size() == 1) c.mode = to_i(d[0]);
            else if (k == "maxAmplitude" && d.size() == 1) c.max_amp = to_d(d[0]);
            else if (k == "position" && d.size() == 1) c.position = d[0];
            else if (k == "fields" && d.size() == 1) c.fields.push_back(d[0]);
            else if (k == "include" && d.size() == 1) c.include.push_back(d[0]);
            else if (k == "writeGradient" && d.size() == 1) c.write_grad = (d[0] == "yes");
            else if (k == "timeStep" && d.size() == 1) c.time_step = to_d(d[0]);
            else if (k == "newmarkBeta" && d.size() == 1) c.beta = to_d(d[0]);
            else if (k == "newmarkGamma" && d.size() == 1) c.gamma = to_d(d[0]);
            else if (k == "simPeriod" && d.size() == 1) c.sim_period = to_d(d[0]);
            else if (k == "rayDampCM" && d.size() == 1) c.ray_cm = to_d(d[0]);
            else if (k == "rayDampCK" && d.size() == 1) c.ray_ck = to_d(d[0]);
            else if (k == "saveSolnHist" && d.size() == 1) c.save_hist = d[0].find("yes") != std::string::npos;
            else if (k == "solnHistDir" && d.size() == 1) c.file = d[0];
            else if (k == "lumpMass" && d.size() == 1) c.lump_mass = d[0].find("yes") != std::string::npos;
            else if (k == "staticLoadTime" && !d.empty())
                for (const auto& x : d) c.load_times.push_back(to_d(x));
            else if (k == "timeSteps" && d.size() == 1 && d[0] != "all" && d[0] != "last") {
                // "all" and "last" are read by the reference but never used
                c.time_steps.push_back(to_i(d[0]));
            } else if (k == "timeSteps" && d.size() > 1) {
                const long a = to_i(d[0]), b = to_i(d[1]), st = d.size() == 3 ? to_i(d[2]) : 1;
                for (long i = a; i < b; i += st) c.time_steps.push_back(i);
            }
        }
    }

    void read_model(const std::string& path) {
        LineReader lr;
        std::vector<std::string> d;
        int cur_type = 0;
        std::vector<NamedSet>* cur_sets = nullptr;
        bool in_mat = false;
        for (const auto& line : read_lines(path, "Model input")) {
            const bool upd = lr.read(line, d);
            const std::string* h = lr.head;
            if (h[0] == "nodes" && d.size() == 4) {
                const long n = to_i(d[0]);
                if (n < 0) die("negative node label");
                if ((size_t)(3 * n + 3) > xyz.size()) xyz.resize(3 * n + 3, 0.0);
                for (int c = 0; c < 3; ++c) xyz[3 * n + c] = to_d(d[1 + c]);
            } else if (h[0] == "elements") {
                if (h[1] == "type" && d.size() == 1) {
                    if (d[0] == "brick8") cur_type = ASEND_BRICK8;
                    else if (d[0] == "brickIM") cur_type = ASEND_BRICKIM;
                    else if (d[0] == "shell4") cur_type = ASEND_SHELL4;
                    else die("element type %s is not supported by this solver", d[0].c_str());
                } else if (h[1] == "connectivity" && d.size() > 1) {
                    const long e = to_i(d[0]);
                    if (e < 0) die("negative element label");
                    if ((size_t)e >= conn.size()) {
                        conn.resize(e + 1);
                        el_type.resize(e + 1, 0);
                    }
                    el_type[e] = cur_type;
                    conn[e].clear();
                    for (size_t i = 1; i < d.size(); ++i) conn[e].push_back((int32_t)to_i(d[i]));
                }
            } else if (h[0] == "sets" && (h[1] == "node" || h[1] == "element") && !h[2].empty()) {
                auto& sets = (h[1] == "node") ? nsets : esets;
                if (upd) {
                    sets.push_back(NamedSet());
                    sets.back().name = h[2];
                    cur_sets = &sets;
                    continue;
                }
                if (cur_sets != &sets || sets.empty()) continue;
                if (d.size() == 1) {
                    sets.back().labels.push_back((int32_t)to_i(d[0]));
                } else if (d.size() > 1) {
                    const long a = to_i(d[0]), b = to_i(d[1]), st = d.size() == 3 ? to_i(d[2]) : 1;
                    for (long i = a; i < b; i += st) sets.back().labels.push_back((int32_t)i);
                }
            } else if (h[0] == "sections") {
                if (h[1] == "type" && d.size() == 1) {
                    sections.push_back(Section());
                    sections.back().type = d[0];
                } else if (sections.empty()) {
                    continue;
                } else if (h[1] == "material" && d.size() == 1) {
                    sections.back().mat = d[0];
                } else if (h[1] == "elementSet" && d.size() == 1) {
                    sections.back().el_set = d[0];
                } else if (h[1] == "orientation" && d.size() == 6) {
                    // x' = xDir normalised, z' = x' x xyVec normalised, y' = z' x x'
                    double* o = sections.back().ori;
                    double v[6];
                    for (int i = 0; i < 6; ++i) v[i] = to_d(d[i]);
                    double mg = 1.0 / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                    for (int i = 0; i < 3; ++i) o[i] = mg * v[i];
                    double z[3] = {o[1] * v[5] - o[2] * v[4], o[2] * v[3] - o[0] * v[5], o[0] * v[4] - o[1] * v[3]};
                    mg = 1.0 / std::sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
                    for (int i = 0; i < 3; ++i) o[6 + i] = mg * z[i];
                    o[3] = o[7] * o[2] - o[8] * o[1];
                    o[4] = o[8] * o[0] - o[6] * o[2];
                    o[5] = o[6] * o[1] - o[7] * o[0];
                    bool id = true;
                    for (int i = 0; i < 9; ++i) id = id && std::fabs(o[i] - ((i % 4 == 0) ? 1.0 : 0.0)) <= 1e-15;
                    sections.back().oriented = !id;
                } else if (h[1] == "layup") {
                    if (h[2] == "zOffset" && d.size() == 1) sections.back().z_offset = to_d(d[0]);
                    else if (h[2] == "layers" && h[3] == "material" && d.size() == 1) {
                        sections.back().lay_mat.push_back(d[0]);
                        sections.back().lay_t.push_back(0.0);
                    } else if (h[2] == "layers" && h[3] == "thickness" && d.size() == 1 &&
                               !sections.back().lay_t.empty()) {
                        sections.back().lay_t.back() = to_d(d[0]);
                    }
                }
            } else if (h[0] == "materials" && !h[1].empty()) {
                if (h[2].empty() && upd) {
                    materials.push_back(Material());
                    materials.back().name = h[1];
                    in_mat = true;
                } else if (in_mat) {
                    Material* cur_mat = &materials.back();
                    if (h[2] == "density" && d.size() == 1) cur_mat->density = to_d(d[0]);
                    else if (h[2] == "elastic" && (h[3] == "E" || h[3] == "nu" || h[3] == "G") && !d.empty()) {
                        double* v = (h[3] == "E") ? cur_mat->E : (h[3] == "nu") ? cur_mat->nu : cur_mat->G;
                        v[0] = to_d(d[0]);
                        v[1] = d.size() == 3 ? to_d(d[1]) : v[0];
                        v[2] = d.size() == 3 ? to_d(d[2]) : v[0];
                    } else if (h[2] == "elastic" && h[3] == "stiffness" && d.size() == 3) {
                        const long r = to_i(d[0]) - 1, c = to_i(d[1]) - 1;
                        if (r < 0 || r > 5 || c < 0 || c > 5) die("stiffness entry out of range");
                        cur_mat->stiff[6 * r + c] = cur_mat->stiff[6 * c + r] = to_d(d[2]);
                        cur_mat->stiffness = true;
                    } else if (h[2] == "damping" && d.size() == 3) {
                        const long r = to_i(d[0]) - 1, c = to_i(d[1]) - 1;
                        if (r < 0 || r > 5 || c < 0 || c > 5) die("damping entry out of range");
                        cur_mat->damping[6 * r + c] = cur_mat->damping[6 * c + r] = to_d(d[2]);
                        cur_mat->damped = true;
                    } else if (h[2] == "thermal" && h[3] == "specHeat" && d.size() == 1) {
                        cur_mat->spec_heat = to_d(d[0]);
                    } else if (h[2] == "thermal" && h[3] == "conductivity" && d.size() == 6) {
                        for (int i = 0; i < 6; ++i) cur_mat->cond[i] = to_d(d[i]);
                    } else if (h[2] == "thermal" && h[3] == "expansion" && d.size() == 6) {
                        for (int i = 0; i < 6; ++i) {
                            cur_mat->expansion[i] = to_d(d[i]);
                            cur_mat->expands = cur_mat->expands || cur_mat->expansion[i] != 0.0;
                        }
                    }
                }
            }
        }
        // individual sets named by label and the "all" sets
        // (read_model_input, S/model_input.rs:804-860)
        const size_t nn = xyz.size() / 3, ne = conn.size();
        for (size_t i = 0; i < nn; ++i) nsets.push_back(NamedSet{std::to_string(i), {(int32_t)i}});
        NamedSet all_n{"all", {}};
        for (size_t i = 0; i < nn; ++i) all_n.labels.push_back((int32_t)i);
        nsets.push_back(all_n);
        for (size_t i = 0; i < ne; ++i) esets.push_back(NamedSet{std::to_string(i), {(int32_t)i}});
        NamedSet all_e{"all", {}};
        for (size_t i = 0; i < ne; ++i) all_e.labels.push_back((int32_t)i);
        esets.push_back(all_e);
        ns_map.clear();
        es_map.clear();
        for (size_t i = 0; i < nsets.size(); ++i) ns_map[nsets[i].name] = i;
        for (size_t i = 0; i < esets.size(); ++i) es_map[esets[i].name] = i;
    }

    // a file with constraints replaces the model's (read_constraint_input
    // reallocates its lists; one without any leaves them)
    void read_constraints(const std::string& path) {
        LineReader lr;
        std::vector<std::string> d;
        bool cur_disp = false;
        const auto lines = read_lines(path, "Constraint input");
        {
            LineReader cnt;
            bool any = false;
            for (const auto& line : lines) {
                cnt.read(line, d);
                any = any || (cnt.head[0] == "constraints" && cnt.head[1] == "type" && d.size() == 1);
            }
            if (!any) return;
            constraints.clear();
            tconstraints.clear();
        }
        std::vector<Constraint>* cur = nullptr;
        for (const auto& line : lines) {
            lr.read(line, d);
            const std::string* h = lr.head;
            if (h[0] != "constraints") continue;
            if (h[1] == "type" && d.size() == 1) {
                if (d[0] != "displacement" && d[0] != "temperature" && d[0] != "concentration" && d[0] != "fluid")
                    die("%s is not a valid constraint type", d[0].c_str());
                cur_disp = (d[0] == "displacement");
                cur = cur_disp ? &constraints : (d[0] == "temperature" ? &tconstraints : nullptr);
                if (cur) {
                    cur->push_back(Constraint());
                    cur->back().type = d[0];
                }
            } else if (!cur || cur->empty()) {
                continue;
            } else if (h[1] == "terms") {
                if (h[2] == "nodeSet" && d.size() == 1) {
                    cur->back().terms.push_back(Term());
                    cur->back().terms.back().set = d[0];
                } else if (h[2] == "dof" && d.size() == 1 && !cur->back().terms.empty()) {
                    cur->back().terms.back().dof = (int)to_i(d[0]);
                } else if (h[2] == "coef" && d.size() == 1 && !cur->back().terms.empty()) {
                    cur->back().terms.back().coef = to_d(d[0]);
                }
            } else if (h[1] == "rhs" && d.size() == 1) {
                cur->back().rhs = to_d(d[0]);
            }
        }
    }

    // likewise for loads (read_load_input)
    void read_loads(const std::string& path) {
        LineReader lr;
        std::vector<std::string> d;
        const std::string elastic = "nodalForce bodyForce gravitational centrifugal surfacePressure surfaceTraction";
        const std::string thermal = "nodalHeatGen bodyHeatGen surfaceFlux";  // read_load_input's lists
        std::vector<Load>* cur = nullptr;
        const auto lines = read_lines(path, "Load input");
        {
            LineReader cnt;
            bool any = false;
            for (const auto& line : lines) {
                cnt.read(line, d);
                any = any || (cnt.head[0] == "loads" && cnt.head[1] == "type" && d.size() == 1);
            }
            if (!any) return;
            loads.clear();
            tloads.clear();
        }
        for (const auto& line : lines) {
            lr.read(line, d);
            const std::string* h = lr.head;
            if (h[0] != "loads") continue;
            if (h[1] == "type" && d.size() == 1) {
                cur = elastic.find(d[0]) != std::string::npos   ? &loads
                      : thermal.find(d[0]) != std::string::npos ? &tloads
                                                                : nullptr;
                if (cur) {
                    cur->push_back(Load());
                    cur->back().type = d[0];
                }
            } else if (!cur || cur->empty()) {
                continue;
            } else if (h[1] == "load" && !d.empty()) {
                for (int i = 0; i < 6; ++i) cur->back().load[i] = (i < (int)d.size()) ? to_d(d[i]) : 0.0;
            } else if (h[1] == "nodeSet" && d.size() == 1) {
                cur->back().node_set = d[0];
            } else if (h[1] == "elementSet" && d.size() == 1) {
                cur->back().el_set = d[0];
            } else if (h[1] == "normDir" && d.size() == 3) {
                for (int i = 0; i < 3; ++i) cur->back().norm_dir[i] = to_d(d[i]);
            } else if (h[1] == "normTolerance" && d.size() == 1) {
                cur->back().norm_tol = to_d(d[0]);
            } else if (h[1] == "center" && d.size() == 3) {
                for (int i = 0; i < 3; ++i) cur->back().center[i] = to_d(d[i]);
            } else if (h[1] == "axis" && d.size() == 3) {
                for (int i = 0; i < 3; ++i) cur->back().axis[i] = to_d(d[i]);
            } else if (h[1] == "angularVelocity" && !d.empty()) {
                cur->back().ang_vel = to_d(d[0]);
            } else if (h[1] == "activeTime" && !d.empty()) {
                cur->back().active[0] = to_d(d[0]);
                cur->back().active[1] = d.size() == 2 ? to_d(d[1]) : 1.0e100;
            }
        }
    }

    // a design-variable or objective file replaces the previous set
    void read_dvars(const std::string& path) {
        LineReader lr;
        std::vector<std::string> d;
        dvars.clear();
        for (const auto& line : read_lines(path, "design variable input")) {
            lr.read(line, d);
            const std::string* h = lr.head;
            if (h[0] != "designVariables") continue;
            if (h[1] == "category" && d.size() == 1) {
                dvars.push_back(DesignVar());
                dvars.back().category = d[0];
            } else if (dvars.empty()) {
                continue;
            } else if (h[1] == "elementSet" && d.size() == 1) {
                dvars.back().el_set = d[0];
            } else if (h[1] == "nodeSet" && d.size() == 1) {
                dvars.back().nd_set = d[0];
            } else if (h[1] == "component" && d.size() == 1) {
                dvars.back().component = (int)to_i(d[0]);
            } else if (h[1] == "coefficients" && d.size() == 1) {
                dvars.back().coefs.push_back(to_d(d[0]));
            }
        }
    }

    void read_dvar_values(const std::string& path) {
        LineReader lr;
        std::vector<std::string> d;
        for (const auto& line : read_lines(path, "design variable value input")) {
            lr.read(line, d);
            if (d.size() == 2) {
                const long i = to_i(d[0]);
                if (i < 0 || (size_t)i >= dvars.size()) die("design variable %ld is not defined", i);
                dvars[i].value = to_d(d[1]);
            }
        }
    }

    void read_objective(const std::string& path) {
        LineReader lr;
        std::vector<std::string> d;
        obj.clear();
        for (const auto& line : read_lines(path, "Objective input")) {
            lr.read(line, d);
            const std::string* h = lr.head;
            if (h[0] != "objectiveTerms") continue;
            if (h[1] == "category" && d.size() == 1) {
                obj.push_back(ObjTerm());
                obj.back().category = d[0];
            } else if (obj.empty()) {
                continue;
            } else if (h[1] == "operator" && d.size() == 1) {
                obj.back().optr = d[0];
            } else if (h[1] == "component" && d.size() == 1) {
                obj.back().component = (int)to_i(d[0]);
            } else if (h[1] == "coefficient" && d.size() == 1) {
                obj.back().coef = to_d(d[0]);
            } else if (h[1] == "exponent" && d.size() == 1) {
                obj.back().expnt = to_d(d[0]);
            } else if (h[1] == "elementSet" && d.size() == 1) {
                obj.back().el_set = d[0];
            } else if (h[1] == "nodeSet" && d.size() == 1) {
                obj.back().nd_set = d[0];
            } else if (h[1] == "targetValue" && d.size() == 1) {
                obj.back().tgt.push_back(to_d(d[0]));
            }
        }
    }

    // ------------------------------------------------------------ analysis
    // element material after design variables (get_solid_stiff_dfd0's
    // modulus_dv / poisson_dv / shear_mod_dv, S/element_properties.rs:700-726)
    void element_materials() {
        const size_t ne = conn.size();
        std::vector<int> sect(ne, -1);
        for (size_t s = 0; s < sections.size(); ++s)
            for (int32_t e : element_set(sections[s].el_set)) sect[e] = (int)s;
        std::vector<std::array<double, 9>> par(ne);
        mat_of_el.assign(ne, 0);
        sec_of_el.assign(ne, 0);
        thk.assign(ne, 0.0);
        rho.assign(ne, 0.0);
        kc.assign(6 * ne, 0.0);
        auto mat_of = [&](const std::string& n) -> const Material& {
            for (const auto& m : materials)
                if (m.name == n) return m;
            die("material %s is not defined", n.c_str());
        };
        for (size_t e = 0; e < ne; ++e) {
            if (sect[e] < 0) die("element %zu has no section", e);
            const Section& s = sections[sect[e]];
            const Material* m;
            if (etype == ASEND_SHELL4) {
                if (s.lay_mat.size() != 1 || s.z_offset != 0.0)
                    die("shell4 sections must have one layer centred on the mid-plane");
                m = &mat_of(s.lay_mat[0]);
                thk[e] = s.lay_t[0];
            } else {
                m = &mat_of(s.mat);
            }
            mat_of_el[e] = (int)(m - materials.data());
            sec_of_el[e] = sect[e];
            std::copy(m->cond, m->cond + 6, kc.begin() + 6 * e);
            for (int c = 0; c < 3; ++c) {
                par[e][c] = m->E[c];
                par[e][3 + c] = m->nu[c];
                par[e][6 + c] = m->G[c];
            }
            rho[e] = m->density;
        }
        for (const auto& dv : dvars) {
            int base;
            if (dv.category == "modulus") base = 0;
            else if (dv.category == "poissonRatio") base = 3;
            else if (dv.category == "shearModulus") base = 6;
            else if (dv.category == "density") base = -1;
            else if (dv.category == "nodeCoord") continue;  // applied to the coordinates (node_coords)
            else die("design variables of category %s are not supported", dv.category.c_str());
            if (base < 0) {
                // density = material density + sum of coefficient x value
                // (get_density_dfd0, S/element_properties.rs:1313-1350)
                if (dv.el_set.empty()) die("design variables of category density need an element set");
                const auto& set = element_set(dv.el_set);
                if (dv.coefs.size() >= 2 && dv.coefs.size() != set.size())
                    die("coefficient list provided for a design variable does not match the length of the element "
                        "set");
                for (size_t k = 0; k < set.size(); ++k)
                    rho[set[k]] += (dv.coefs.empty() ? 1.0 : (dv.coefs.size() == 1 ? dv.coefs[0] : dv.coefs[k])) *
                                   dv.value;
                continue;
            }
            if (dv.component < 1 || dv.component > 3) die("design variable component out of range");
            if (dv.el_set.empty()) die("design variables of category %s need an element set", dv.category.c_str());
            const auto& set = element_set(dv.el_set);
            if (dv.coefs.size() >= 2 && dv.coefs.size() != set.size())
                die("coefficient list provided for a design variable does not match the length of the element set");
            for (size_t k = 0; k < set.size(); ++k) {
                const double c = dv.coefs.empty() ? 1.0 : (dv.coefs.size() == 1 ? dv.coefs[0] : dv.coefs[k]);
                par[set[k]][base + dv.component - 1] += c * dv.value;
            }
        }
        E.assign(ne, 0.0);
        nu.assign(ne, 0.0);
        cmats.clear();
        bool material_dvs = false;
        for (const auto& dv : dvars)
            material_dvs = material_dvs || dv.category == "modulus" || dv.category == "poissonRatio" ||
                           dv.category == "shearModulus";
        for (size_t e = 0; e < ne; ++e) {
            const auto& p = par[e];
            const Material& me = materials[mat_of_el[e]];
            if (me.stiffness || !isotropic(p)) {
                // general material (get_solid_stiff_dfd0): the stiffness entries, or
                // the inverse of the compliance of the nine engineering constants
                if (etype != ASEND_BRICK8) die("element %zu: anisotropic materials are supported for brick8 models", e);
                if (material_dvs) die("material design variables on anisotropic materials are not supported");
                if (cmats.empty()) cmats.assign(36 * sections.size(), 0.0);
                double* C = cmats.data() + 36 * sec_of_el[e];
                if (me.stiffness) std::copy(me.stiff, me.stiff + 36, C);
                else compliance_inverse(p, C);
                rotate_to_global(sections[sec_of_el[e]], C);
                if (me.expands && sections[sec_of_el[e]].oriented)
                    die("thermal expansion in an oriented section is not supported by this solver");
                E[e] = 1.0;
                nu[e] = 0.3;
                continue;
            }
            const double Ee = p[0], ne_ = p[3];
            const double G = Ee / (2.0 * (1.0 + ne_));
            (void)G;
            E[e] = Ee;
            nu[e] = ne_;
        }
        if (!cmats.empty()) {
            // an element of an isotropic material in a model with anisotropic
            // ones takes its material matrix too
            for (size_t e = 0; e < ne; ++e)
                if (E[e] != 1.0 || nu[e] != 0.3 || cmats[36 * sec_of_el[e]] == 0.0) {
                    double* C = cmats.data() + 36 * sec_of_el[e];
                    compliance_inverse(par[e], C);
                    E[e] = 1.0;
                    nu[e] = 0.3;
                }
        }
    }

    bool isotropic(const std::array<double, 9>& p) const {
        const double Ee = p[0], ne_ = p[3];
        const double G = Ee / (2.0 * (1.0 + ne_));
        auto same = [](double a, double b) { return std::fabs(a - b) <= 1e-12 * std::fabs(b); };
        return same(p[1], Ee) && same(p[2], Ee) && same(p[4], ne_) && same(p[5], ne_) &&
               (etype == ASEND_SHELL4 ? same(p[6], G) : (same(p[6], G) && same(p[7], G) && same(p[8], G)));
    }

    // A section's material matrix in the global frame: the element works in
    // the local axes R (rows), strain_local = T strain_global (engineering
    // Voigt), so C_global = T^T C_local T.
    static void rotate_to_global(const Section& sec, double* C) {
        if (!sec.oriented) return;
        const double* R = sec.ori;
        const int pr[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
        double T[6][6];
        for (int k = 0; k < 6; ++k) {
            double eg[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // tensor of engineering basis strain k
            const int a = pr[k][0], b = pr[k][1];
            if (k < 3) eg[a][a] = 1.0;
            else eg[a][b] = eg[b][a] = 0.5;
            double el[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double v = 0.0;
                    for (int p = 0; p < 3; ++p)
                        for (int q = 0; q < 3; ++q) v += R[3 * i + p] * eg[p][q] * R[3 * j + q];
                    el[i][j] = v;
                }
            for (int r = 0; r < 6; ++r) T[r][k] = r < 3 ? el[pr[r][0]][pr[r][1]] : 2.0 * el[pr[r][0]][pr[r][1]];
        }
        double Cl[36], Cg[36];
        std::copy(C, C + 36, Cl);
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) {
                double v = 0.0;
                for (int p = 0; p < 6; ++p)
                    for (int q = 0; q < 6; ++q) v += T[p][i] * Cl[6 * p + q] * T[q][j];
                Cg[6 * i + j] = v;
            }
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) C[6 * i + j] = 0.5 * (Cg[6 * i + j] + Cg[6 * j + i]);
    }

    // C = S^-1, S the compliance of E1..3, nu12, nu13, nu23, G12, G13, G23
    // (get_solid_stiff_dfd0, S/element_properties.rs:728-761), by Gauss-Jordan
    // elimination with partial pivoting
    static void compliance_inverse(const std::array<double, 9>& p, double* C) {
        double S[6][12] = {};
        S[0][0] = 1.0 / p[0];
        S[0][1] = S[1][0] = -p[3] / p[0];
        S[0][2] = S[2][0] = -p[4] / p[0];
        S[1][1] = 1.0 / p[1];
        S[1][2] = S[2][1] = -p[5] / p[1];
        S[2][2] = 1.0 / p[2];
        S[3][3] = 1.0 / p[6];
        S[4][4] = 1.0 / p[7];
        S[5][5] = 1.0 / p[8];
        for (int i = 0; i < 6; ++i) S[i][6 + i] = 1.0;
        for (int c = 0; c < 6; ++c) {
            int piv = c;
            for (int r = c + 1; r < 6; ++r)
                if (std::fabs(S[r][c]) > std::fabs(S[piv][c])) piv = r;
            if (!(std::fabs(S[piv][c]) > 0.0)) die("singular material compliance");
            if (piv != c)
                for (int k = 0; k < 12; ++k) std::swap(S[c][k], S[piv][k]);
            const double ip = 1.0 / S[c][c];
            for (int k = 0; k < 12; ++k) S[c][k] *= ip;
            for (int r = 0; r < 6; ++r) {
                if (r == c) continue;
                const double f = S[r][c];
                for (int k = 0; k < 12; ++k) S[r][k] -= f * S[c][k];
            }
        }
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) C[6 * i + j] = 0.5 * (S[i][6 + j] + S[j][6 + i]);
    }

    // coordinates after nodeCoord design variables: initial + sum of
    // coefficient x value (Node::get_crd_dfd0, S/node_meth.rs:255-282); the
    // reordering keeps the initial coordinates
    std::vector<double> xyz_dv;
    void node_coords() {
        xyz_dv = xyz;
        for (const auto& dv : dvars) {
            if (dv.category != "nodeCoord") continue;
            if (dv.component < 1 || dv.component > 3) die("design variable component out of range");
            if (dv.nd_set.empty()) die("design variables of category nodeCoord need a node set");
            const auto& set = node_set(dv.nd_set, false);
            if (dv.coefs.size() >= 2 && dv.coefs.size() != set.size())
                die("coefficient list provided for a design variable does not match the length of the node set");
            for (size_t k = 0; k < set.size(); ++k)
                xyz_dv[3 * (size_t)set[k] + dv.component - 1] +=
                    (dv.coefs.empty() ? 1.0 : (dv.coefs.size() == 1 ? dv.coefs[0] : dv.coefs[k])) * dv.value;
        }
    }

    void analysis_prep(const Command& c) {
        const size_t nn = xyz.size() / 3, ne = conn.size();
        if (nn == 0 || ne == 0) die("the model has no nodes or no elements");
        etype = el_type[0];
        for (size_t e = 0; e < ne; ++e)
            if (el_type[e] != etype) die("models with more than one element type are not supported");
        const int npe = (etype == ASEND_SHELL4) ? 4 : 8;
        dpn = (etype == ASEND_SHELL4) ? 6 : 3;
        std::vector<int32_t> cflat(npe * ne);
        for (size_t e = 0; e < ne; ++e) {
            if ((int)conn[e].size() != npe) die("element %zu has %zu nodes", e, conn[e].size());
            for (int a = 0; a < npe; ++a) {
                if (conn[e][a] < 0 || (size_t)conn[e][a] >= nn) die("element %zu names an undefined node", e);
                cflat[npe * e + a] = conn[e][a];
            }
        }
        // constraint structure for the ordering (add_constraint_conn)
        std::vector<int64_t> cptr{0}, tptr{0};
        std::vector<int32_t> setn;
        // the displacement constraints, then the temperature ones (reorder_nodes, S/model_prep.rs:98-99)
        for (const auto* list : {&constraints, &tconstraints})
            for (const auto& k : *list) {
                for (const auto& t : k.terms) {
                    for (int32_t n : node_set(t.set, false)) setn.push_back(n);
                    tptr.push_back((int64_t)setn.size());
                }
                cptr.push_back((int64_t)tptr.size() - 1);
            }
        int64_t bd;
        if (c.method == "direct" || c.method.empty()) bd = 2000000000;
        // an unset solverBlockDim (2e9) takes the automatic size (S/model_prep.rs:892-901)
        else bd = (c.block_dim > 0 && c.block_dim != 2000000000) ? c.block_dim : asend_block_dim_auto((int64_t)nn);
        solver_block_dim = bd;
        std::vector<int32_t> ordered(nn);
        dof_index.assign(6 * nn, -1);
        int rc = asend_reorder_nodes((int64_t)nn, xyz.data(), (int64_t)ne, npe, dpn, cflat.data(),
                                     (int64_t)(cptr.size() - 1), cptr.data(), tptr.data(), setn.data(), bd,
                                     ordered.data(), dof_index.data(), &n_dof);
        if (rc != ASEND_OK) die("node reordering failed (%d)", rc);
        // penalty DOFs: one-term clamps (Constraint::build_mat)
        pen.clear();
        for (const auto& k : constraints) {
            if (k.terms.size() != 1 || std::fabs(std::fabs(k.terms[0].coef) - 1.0) > 0.0 || k.rhs != 0.0)
                die("only single-term displacement constraints with coefficient 1 and rhs 0 are supported");
            const int dof = k.terms[0].dof;
            if (dof < 1 || dof > dpn) die("constraint dof %d out of range", dof);
            for (int32_t n : node_set(k.terms[0].set, false)) pen.push_back(dof_index[6 * n + dof - 1]);
        }
        element_materials();
        node_coords();
        if (c.thermal) thermal_prep(cflat);
        if (!c.elastic) return;
        if (!ctx) check(asend_ctx_create(0, &ctx), "asend_ctx_create");
        check(asend_upload_mesh(ctx, etype, (int64_t)nn, xyz_dv.data(), dof_index.data(), n_dof, (int64_t)ne,
                                cflat.data(), E.data(), nu.data(), rho.data()),
              "asend_upload_mesh");
        if (etype == ASEND_SHELL4) check(asend_set_thickness(ctx, thk.data()), "asend_set_thickness");
        if (!cmats.empty()) {
            std::vector<int32_t> idx(sec_of_el.begin(), sec_of_el.end());
            check(asend_set_material_stiffness(ctx, (int64_t)sections.size(), cmats.data(), idx.data()),
                  "asend_set_material_stiffness");
        }
        check(asend_set_penalty(ctx, (int64_t)pen.size(), pen.data(), 0.0), "asend_set_penalty");
        // applied loads (build_elastic_app_load): one vector per load, summed
        // over the loads active at a time by loads_at()
        for (auto& l : loads) {
            l.vec.assign(n_dof, 0.0);
            double* lf = l.vec.data();
            if (l.type == "nodalForce") {
                for (int32_t n : node_set(l.node_set, false))
                    for (int c = 0; c < dpn; ++c) lf[dof_index[6 * n + c]] += l.load[c];
            } else if (l.type == "surfaceTraction" || l.type == "surfacePressure") {
                if (etype == ASEND_SHELL4) die("%s on shell elements is not supported", l.type.c_str());
                const auto& set = element_set(l.el_set);
                if (l.type == "surfaceTraction")
                    rc = asend_surface_traction(etype, (int64_t)nn, xyz_dv.data(), (int64_t)ne, cflat.data(),
                                                dof_index.data(), (int64_t)set.size(), set.data(), l.load,
                                                l.norm_dir, l.norm_tol, n_dof, lf);
                else
                    rc = asend_surface_pressure(etype, (int64_t)nn, xyz_dv.data(), (int64_t)ne, cflat.data(),
                                                dof_index.data(), (int64_t)set.size(), set.data(), l.load[0],
                                                l.norm_dir, l.norm_tol, n_dof, lf);
                if (rc != ASEND_OK) die("%s load failed (%d)", l.type.c_str(), rc);
            } else if (l.type == "bodyForce" || l.type == "gravitational" || l.type == "centrifugal") {
                if (etype == ASEND_SHELL4) die("%s on shell elements is not supported", l.type.c_str());
                const auto& set = element_set(l.el_set);
                const int kind = l.type == "bodyForce" ? 0 : l.type == "gravitational" ? 1 : 2;
                rc = asend_body_load(etype, kind, (int64_t)nn, xyz_dv.data(), (int64_t)ne, cflat.data(), dof_index.data(),
                                     (int64_t)set.size(), set.data(), rho.data(), l.load, l.center, l.axis, l.ang_vel,
                                     n_dof, lf);
                if (rc != ASEND_OK) die("%s load failed (%d)", l.type.c_str(), rc);
            } else {
                die("loads of type %s are not supported", l.type.c_str());
            }
        }
        loads_at(0.0, f);
    }

    // Heat conduction set-up (the thermal half of analysis_prep): a second
    // device context over the same DOF numbering with the elements'
    // conductivities, the temperature penalties (single-term constraints,
    // coefficient +-1, any rhs: Constraint::build_mat / get_load,
    // S/const_meth.rs:30-119) and one load vector per thermal load
    // (build_thermal_app_load, S/model_analysis.rs:77-127).
    void thermal_prep(const std::vector<int32_t>& cflat) {
        if (etype != ASEND_BRICK8) die("thermal solves are supported for brick8 models");
        const size_t nn = xyz.size() / 3, ne = conn.size();
        tpen.clear();
        tpen3.clear();
        tpen_rhs.clear();
        for (const auto& k : tconstraints) {
            if (k.terms.size() != 1 || std::fabs(std::fabs(k.terms[0].coef) - 1.0) > 0.0)
                die("only single-term temperature constraints with coefficient +-1 are supported");
            for (int32_t n : node_set(k.terms[0].set, false)) {
                tpen.push_back(dof_index[6 * n]);
                tpen_rhs.push_back(k.terms[0].coef * k.rhs);
                for (int c = 0; c < 3; ++c) tpen3.push_back(dof_index[6 * n + c]);
            }
        }
        if (!tctx) check_t(asend_ctx_create(0, &tctx), "asend_ctx_create");
        const std::vector<double> one(ne, 1.0), nu0(ne, 0.3);
        // heat capacity per element: density x specific heat (get_rtm_dfd0 with actual_props)
        std::vector<double> rc(ne, 0.0);
        for (size_t e = 0; e < ne; ++e) rc[e] = rho[e] * materials[mat_of_el[e]].spec_heat;
        check_t(asend_upload_mesh(tctx, etype, (int64_t)nn, xyz_dv.data(), dof_index.data(), n_dof, (int64_t)ne,
                                  cflat.data(), one.data(), nu0.data(), rc.data()),
                "asend_upload_mesh");
        check_t(asend_set_conductivity(tctx, kc.data()), "asend_set_conductivity");
        // the temperature's two companion DOFs are penalised at the same nodes
        check_t(asend_set_penalty(tctx, (int64_t)tpen3.size(), tpen3.data(), 0.0), "asend_set_penalty");
        // a transient job's first matrix, which fixes the penalty scale, holds
        // the capacity term (build_thermal_soln_load(true) at analysis_prep)
        const double c1 = solve_cmd.dynamic ? 1.0 / (solve_cmd.gamma * solve_cmd.time_step) : 0.0;
        check_t(asend_assemble(tctx, c1), "asend_assemble");
        if (solve_cmd.dynamic) {
            if (!cctx) check_t(asend_ctx_create(0, &cctx), "asend_ctx_create");
            const std::vector<double> zk(6 * ne, 0.0);
            check_t(asend_upload_mesh(cctx, etype, (int64_t)nn, xyz_dv.data(), dof_index.data(), n_dof, (int64_t)ne,
                                      cflat.data(), one.data(), nu0.data(), rc.data()),
                    "asend_upload_mesh");
            check_t(asend_set_conductivity(cctx, zk.data()), "asend_set_conductivity");
            check_t(asend_assemble(cctx, 1.0), "asend_assemble");
        }
        for (auto& l : tloads) {
            l.vec.assign(n_dof, 0.0);
            double* lf = l.vec.data();
            int rc = ASEND_OK;
            if (l.type == "nodalHeatGen") {
                for (int32_t n : node_set(l.node_set, false)) lf[dof_index[6 * n]] += l.load[0];
            } else if (l.type == "bodyHeatGen") {
                const auto& set = element_set(l.el_set);
                const double q[3] = {l.load[0], 0.0, 0.0};
                const double zero[3] = {0.0, 0.0, 0.0};
                rc = asend_body_load(etype, 0, (int64_t)nn, xyz_dv.data(), (int64_t)ne, cflat.data(),
                                     dof_index.data(), (int64_t)set.size(), set.data(), nullptr, q, zero, zero, 0.0,
                                     n_dof, lf);
            } else {  // surfaceFlux
                const auto& set = element_set(l.el_set);
                rc = asend_surface_flux(etype, (int64_t)nn, xyz_dv.data(), (int64_t)ne, cflat.data(),
                                        dof_index.data(), (int64_t)set.size(), set.data(), l.load[0], l.norm_dir,
                                        l.norm_tol, n_dof, lf);
            }
            if (rc != ASEND_OK) die("%s load failed (%d)", l.type.c_str(), rc);
        }
        t3.assign(n_dof, 0.0);
        t3dot.assign(n_dof, 0.0);
        temp.assign(nn, 0.0);
        tdot.assign(nn, 0.0);
    }

    // The thermal Newmark loop of Model::solve (S/model_analysis.rs:596-603,
    // solve_step :331-370, Node::update_tdot / advance_temp,
    // S/node_meth.rs:140-171): per step at time t, with the predicted rate
    // p = -((1 - gamma) / gamma) Tdot, (K + C / (gamma dt) + sf P) dT =
    // Q(t) - K T - C p + sf (c rhs - T), T += dT,
    // Tdot = (dT / dt - (1 - gamma) Tdot) / gamma.  K_eff T comes from the
    // conduction context, C x from the capacity context.
    void solve_thermal_dynamic() {
        const Command& c = solve_cmd;
        if (!(c.time_step > 0.0)) die("dynamic solves need timeStep > 0");
        const double g = c.gamma, dt = c.time_step, c1 = 1.0 / (g * dt);
        const size_t nn = xyz.size() / 3;
        if (c.save_hist) write_state(0, u, v, a);
        std::vector<double> b(n_dof), ke(n_dof), cx(n_dof), x(n_dof), dT;
        const double sf = asend_penalty_scale(tctx);
        double t = 0.0;
        long step = 1;
        while (t < c.sim_period) {
            std::fill(b.begin(), b.end(), 0.0);
            for (const auto& l : tloads)
                if (t >= l.active[0] && t <= l.active[1])
                    for (int64_t i = 0; i < n_dof; ++i) b[i] += l.vec[i];
            check_t(asend_spmv(tctx, 64, t3.data(), ke.data()), "asend_spmv");  // (K + c1 C + sf P) T
            for (int64_t i = 0; i < n_dof; ++i) x[i] = c1 * t3[i] + ((1.0 - g) / g) * t3dot[i];
            check_t(asend_spmv(cctx, 64, x.data(), cx.data()), "asend_spmv");
            for (int64_t i = 0; i < n_dof; ++i) b[i] += cx[i] - ke[i];
            for (size_t k = 0; k < tpen.size(); ++k) b[tpen[k]] += sf * tpen_rhs[k];
            device_solve(tctx, tldl_ok, tbw_printed, 3 * solver_block_dim, 3, b, dT, false);
            for (int64_t i = 0; i < n_dof; ++i) {
                t3dot[i] = (dT[i] / dt - (1.0 - g) * t3dot[i]) / g;
                t3[i] += dT[i];
            }
            for (size_t n = 0; n < nn; ++n) {
                const int64_t d = dof_index[6 * n];
                temp[n] = d >= 0 ? t3[d] : 0.0;
                tdot[n] = d >= 0 ? t3dot[d] : 0.0;
            }
            thermal_solved = true;
            if (c.save_hist) write_state(step++, u, v, a);
            t += dt;
        }
        std::printf("device transient heat conduction: %ld steps\n", step - 1);
    }

    // The thermal branch of solve_step (S/model_analysis.rs:331-370): the
    // load of the thermal loads active at t, minus the element flux of the
    // current temperatures, plus the penalty load sf (c rhs - T); solved for
    // the increment, which is added to the temperatures.
    void thermal_step(double t) {
        std::vector<double> b(n_dof, 0.0), kt(n_dof, 0.0), dt;
        for (const auto& l : tloads)
            if (t >= l.active[0] && t <= l.active[1])
                for (int64_t i = 0; i < n_dof; ++i) b[i] += l.vec[i];
        check_t(asend_spmv(tctx, 64, t3.data(), kt.data()), "asend_spmv");  // K T + sf P T
        const double sf = asend_penalty_scale(tctx);
        for (int64_t i = 0; i < n_dof; ++i) b[i] -= kt[i];
        for (size_t k = 0; k < tpen.size(); ++k) b[tpen[k]] += sf * tpen_rhs[k];
        device_solve(tctx, tldl_ok, tbw_printed, 3 * solver_block_dim, 3, b, dt, true);
        for (int64_t i = 0; i < n_dof; ++i) t3[i] += dt[i];
        const size_t nn = xyz.size() / 3;
        for (size_t n = 0; n < nn; ++n) temp[n] = dof_index[6 * n] >= 0 ? t3[dof_index[6 * n]] : 0.0;
        thermal_solved = true;
    }

    // the expansion load of the current temperatures (get_ruk_dfd0's thermal
    // strain, S/element_equations.rs:272-298) when a thermal solve ran and a
    // material expands; zero otherwise
    std::vector<double> fth;
    void expansion_load() {
        fth.assign(n_dof, 0.0);
        if (!solve_cmd.thermal || !thermal_solved) return;
        bool any = false;
        for (const auto& m : materials) any = any || m.expands;
        if (!any) return;
        const size_t ne = conn.size(), nn = xyz.size() / 3;
        std::vector<double> al(6 * ne);
        for (size_t e = 0; e < ne; ++e)
            std::copy(materials[mat_of_el[e]].expansion, materials[mat_of_el[e]].expansion + 6, al.begin() + 6 * e);
        std::vector<int32_t> cflat(8 * ne), cidx(sec_of_el.begin(), sec_of_el.end());
        for (size_t e = 0; e < ne; ++e)
            for (int k = 0; k < 8; ++k) cflat[8 * e + k] = conn[e][k];
        const int rc = asend_thermal_strain_load(etype, (int64_t)nn, xyz_dv.data(), (int64_t)ne, cflat.data(),
                                                 dof_index.data(), E.data(), nu.data(),
                                                 cmats.empty() ? nullptr : cmats.data(),
                                                 cmats.empty() ? nullptr : cidx.data(), al.data(), temp.data(), n_dof,
                                                 fth.data());
        if (rc != ASEND_OK) die("thermal expansion load failed (%d)", rc);
    }

    // reactionHeatGen: the element flux of the temperatures without the
    // penalty (build_thermal_soln_load(false) negated, S/model_output.rs:201-208)
    std::vector<double> reaction_heat() {
        if (!thermal_solved) die("reactionHeatGen needs a thermal solution");
        std::vector<double> r(n_dof, 0.0);
        check_t(asend_spmv(tctx, 64, t3.data(), r.data()), "asend_spmv");
        const double sf = asend_penalty_scale(tctx);
        for (int64_t d : tpen3) r[d] -= sf * t3[d];
        if (solve_cmd.dynamic) {  // K T + C Tdot: remove c1 C T, add C Tdot
            const double c1 = 1.0 / (solve_cmd.gamma * solve_cmd.time_step);
            std::vector<double> x(n_dof), cx(n_dof);
            for (int64_t i = 0; i < n_dof; ++i) x[i] = t3dot[i] - c1 * t3[i];
            check_t(asend_spmv(cctx, 64, x.data(), cx.data()), "asend_spmv");
            for (int64_t i = 0; i < n_dof; ++i) r[i] += cx[i];
        }
        return r;
    }

    // sum of the loads whose activeTime holds t (build_elastic_app_load's
    // time >= active_time[0] && time <= active_time[1])
    void loads_at(double t, std::vector<double>& out) const {
        out.assign(n_dof, 0.0);
        for (const auto& l : loads)
            if (t >= l.active[0] && t <= l.active[1])
                for (int64_t i = 0; i < n_dof; ++i) out[i] += l.vec[i];
    }

    // ------------------------------------------------------------ state files
    // solnTStep<step>.out in the history directory (write_time_step_soln,
    // S/model_output.rs:14-82): per node label, its num_dof displacements,
    // velocities and accelerations as big-endian doubles; elements with
    // internal DOFs would add theirs (brickIM, refused here).
    std::string state_path(long step) const {
        return solve_cmd.file + "/solnTStep" + std::to_string(step) + ".out";
    }

    static void put_be(std::FILE* fp, double v) {
        uint64_t b;
        std::memcpy(&b, &v, 8);
        unsigned char o[8];
        for (int i = 0; i < 8; ++i) o[i] = (unsigned char)(b >> (8 * (7 - i)));
        std::fwrite(o, 1, 8, fp);
    }

    static bool get_be(std::FILE* fp, double* v) {
        unsigned char o[8];
        if (std::fread(o, 1, 8, fp) != 8) return false;
        uint64_t b = 0;
        for (int i = 0; i < 8; ++i) b = (b << 8) | o[i];
        std::memcpy(v, &b, 8);
        return true;
    }

    void write_state(long step, const std::vector<double>& su, const std::vector<double>& sv,
                     const std::vector<double>& sa) {
        const std::string path = state_path(step);
        std::FILE* fp = std::fopen(path.c_str(), "wb");
        if (!fp) die("couldn't open file, %s", path.c_str());
        const size_t nn = xyz.size() / 3;
        for (size_t n = 0; n < nn; ++n) {
            if (solve_cmd.thermal) {  // temperature and its rate
                put_be(fp, temp[n]);
                put_be(fp, tdot[n]);
            }
            if (!solve_cmd.elastic) continue;
            for (const auto* vec : {&su, &sv, &sa})
                for (int c = 0; c < dpn; ++c) {
                    const int64_t d = dof_index[6 * n + c];
                    put_be(fp, d >= 0 ? (*vec)[d] : 0.0);
                }
        }
        if (solve_cmd.elastic && etype == ASEND_BRICKIM) {
            // the elements' internal DOFs after the nodes; a dynamic solve
            // keeps none (brickIM Newmark runs on the condensed operator)
            if (solve_cmd.dynamic) die("solution history of dynamic brickIM solves (internal DOFs) is not supported");
            std::vector<double> al(9 * conn.size());
            check(asend_brickim_internal(ctx, solve_cmd.nonlinear ? nullptr : su.data(), al.data()),
                  "asend_brickim_internal");
            for (double x : al) put_be(fp, x);
        }
        std::fclose(fp);
    }

    void read_state(long step, std::vector<double>& su, std::vector<double>& sv, std::vector<double>& sa) {
        const std::string path = state_path(step);
        std::FILE* fp = std::fopen(path.c_str(), "rb");
        if (!fp) die("couldn't open file, %s", path.c_str());
        su.assign(n_dof, 0.0);
        sv.assign(n_dof, 0.0);
        sa.assign(n_dof, 0.0);
        const size_t nn = xyz.size() / 3;
        for (size_t n = 0; n < nn; ++n) {
            if (solve_cmd.thermal) {
                double tn, rate;
                if (!get_be(fp, &tn) || !get_be(fp, &rate)) die("problem reading file, %s", path.c_str());
                temp[n] = tn;
                tdot[n] = rate;
                if (dof_index[6 * n] >= 0) {
                    t3[dof_index[6 * n]] = tn;
                    t3dot[dof_index[6 * n]] = rate;
                }
            }
            if (!solve_cmd.elastic) continue;
            for (auto* vec : {&su, &sv, &sa})
                for (int c = 0; c < dpn; ++c) {
                    double x;
                    if (!get_be(fp, &x)) die("problem reading file, %s", path.c_str());
                    const int64_t d = dof_index[6 * n + c];
                    if (d >= 0) (*vec)[d] = x;
                }
        }
        if (solve_cmd.elastic && etype == ASEND_BRICKIM) {
            std::vector<double> al(9 * conn.size());
            for (double& x : al)
                if (!get_be(fp, &x)) die("problem reading file, %s", path.c_str());
            if (solve_cmd.nonlinear) check(asend_brickim_set_internal(ctx, al.data()), "asend_brickim_set_internal");
        }
        std::fclose(fp);
    }

    // K x = b through the device as solve_step (S/model_analysis.rs:470-501):
    // the direct branch factors K_eff over one block (ldl_factor) and solves
    // with the factor (ldl_solve); the iterative branch runs conj_grad_sparse
    // preconditioned by the factor over blocks of 6 solverBlockDim DOFs,
    // with its absolute stop rule.  The factor is kept for the adjoint solve
    // and redone after a new assembly.  A factor that does not fit the device
    // falls back to FP64 Jacobi PCG (to a 1e-12 relative residual for direct,
    // the same absolute rule for iterative), saying so.
    void device_solve(const std::vector<double>& b, std::vector<double>& x, bool report) {
        device_solve(ctx, ldl_ok, bw_printed, 6 * (int64_t)solver_block_dim, 1, b, x, report);
    }

    // on context c: LDL^T blocks of block_dofs DOFs for the iterative
    // method; maxBandwidth is printed in units of bw_div DOFs (3 for the
    // conduction context, whose temperature DOFs are every third)
    void device_solve(asend_ctx* c, bool& ldl_ok, bool& bw_printed, int64_t block_dofs, int bw_div,
                      const std::vector<double>& b, std::vector<double>& x, bool report) {
        const bool iterative = (solve_cmd.method == "iterative");
        asend_ldl_info li;
        std::memset(&li, 0, sizeof li);
        const int frc = ldl_ok ? asend_ldl_factor(c, iterative ? block_dofs : 0, &li) : -1;
        if (ldl_ok && frc != ASEND_OK) {
            if (frc != ASEND_ERR_UNSUPPORTED) die("asend_ldl_factor failed (%d): %s", frc, asend_last_error(c));
            std::printf("note: LDL^T factor not available on the device (%s); FP64 Jacobi PCG instead\n",
                        asend_last_error(c));
            ldl_ok = false;
        }
        if (ldl_ok && !bw_printed) {
            std::printf("maxBandwidth: %lld\n", (long long)(li.max_bandwidth / bw_div));
            if (!li.pos_def)
                std::printf("Warning: preconditioning matrix for conjugate gradient solver is not positive "
                            "definite.\n");
            bw_printed = true;
        }
        asend_solve_opts o;
        std::memset(&o, 0, sizeof o);
        o.method = ASEND_SOLVE_PCG_FP64;
        o.precond = ldl_ok ? ASEND_PC_LDL : ASEND_PC_JACOBI;
        o.check_every = 1;
        const int64_t max_it = solve_cmd.max_it > 0 ? solve_cmd.max_it : n_dof;  // S/model_prep.rs:237-243
        if (iterative) {
            o.stop_rule = ASEND_STOP_REF_ABS;
            o.tol = solve_cmd.conv_tol;
            o.max_it = max_it;
        } else if (ldl_ok) {
            o.method = ASEND_SOLVE_DIRECT;
        } else {
            o.stop_rule = ASEND_STOP_REL_RESIDUAL;
            o.tol = 1e-12;
            o.max_it = 100 * n_dof + 1000;
            o.check_every = 16;
        }
        asend_solve_stats st;
        std::memset(&st, 0, sizeof st);
        x.assign(n_dof, 0.0);
        if (const int rc = asend_solve(c, b.data(), x.data(), &o, &st))
            die("asend_solve failed (%d): %s", rc, asend_last_error(c));
        if (report && iterative) {
            if (!st.converged)
                std::printf("Warning: Conjugate gradient solver did not converge to specified tolerance, %s within "
                            "tne maximum number of iterations, %lld\n",
                            fmt_display(solve_cmd.conv_tol).c_str(), (long long)max_it);
            std::printf("Total CG iterations: %lld\n", (long long)st.iterations);
            std::printf("Final residual norm: %s\n", fmt_display(st.ref_residual).c_str());
        } else if (report) {
            std::printf("device solve (%s): relative residual %s\n",
                        o.method == ASEND_SOLVE_DIRECT ? "LDL^T factor" : "FP64 Jacobi PCG",
                        fmt_e12(st.rel_residual).c_str());
        }
    }

    // Model::solve (S/model_analysis.rs:560-669), linear elastic: a static
    // solve per staticLoadTime with the loads active at that time, or the
    // Newmark loop over [0, simPeriod) with the loads active at the start of
    // each step; with saveSolnHist the state after each load time / step
    // goes to solnTStep<i>.out (the dynamic loop writes the initial state as
    // step 0 first).
    // Static steps solve for the increment of the state (solve_step builds
    // the load minus the element residual of the current state plus the
    // penalty load, S/model_analysis.rs:331-370,450-534): the thermal
    // branch first, then the elastic one.  Thermal and elastic solves run
    // side by side only without thermal expansion (no coupling term).
    void solve(const Command& c) {
        if (c.nonlinear && (c.dynamic || !c.elastic))
            die("geometric nonlinearity is supported for static elastic solves by this solver");
        if (!c.elastic && !c.thermal) die("the solve has neither the elastic nor the thermal field");
        if (c.thermal && c.dynamic && c.elastic)
            die("transient heat conduction beside a dynamic elastic solve is not supported by this solver");
        if (c.thermal && c.elastic && (c.nonlinear || c.dynamic))
            for (const auto& m : materials)
                if (m.expands)
                    die("thermal expansion is coupled for linear static solves by this solver (not nonlinear or dynamic)");
        solve_cmd = c;
        if (solve_cmd.load_times.empty()) solve_cmd.load_times.push_back(0.0);
        const int64_t n_prev = n_dof;
        if (init_pending && !first_solve_done) {  // the initial state starts the first solve
            nd_state = init_state;
            nd_state.resize(20 * (xyz.size() / 3), 0.0);
            state_kept = true;
        }
        first_solve_done = true;
        analysis_prep(solve_cmd);
        // a state set by setSolnToMode / zeroSolution on the same numbering is
        // the starting state (the reference's solves start from the nodes' state)
        if (!(state_kept && n_prev == n_dof && (int64_t)u.size() == n_dof)) u.assign(n_dof, 0.0);
        if (!(state_kept && (int64_t)v.size() == n_dof)) v.assign(n_dof, 0.0);
        if (!(state_kept && (int64_t)a.size() == n_dof)) a.assign(n_dof, 0.0);
        if (state_kept && !nd_state.empty()) apply_node_state();  // read node results, in this numbering
        nd_state.clear();
        state_kept = false;
        if (c.dynamic && c.thermal) {
            solve_thermal_dynamic();
        } else if (c.dynamic) {
            solve_dynamic();
        } else if (c.nonlinear) {
            solve_nonlinear();
        } else {
            if (c.elastic) check(asend_assemble(ctx, 0.0), "asend_assemble");
            std::vector<double> b(n_dof), ku(n_dof), du;
            for (size_t i = 0; i < solve_cmd.load_times.size(); ++i) {
                if (c.thermal) thermal_step(solve_cmd.load_times[i]);
                if (c.elastic) {
                    loads_at(solve_cmd.load_times[i], f);
                    expansion_load();
                    check(asend_spmv(ctx, 64, u.data(), ku.data()), "asend_spmv");  // K u + sf P u
                    for (int64_t k = 0; k < n_dof; ++k) b[k] = f[k] + fth[k] - ku[k];
                    device_solve(b, du, true);
                    for (int64_t k = 0; k < n_dof; ++k) u[k] += du[k];
                }
                if (solve_cmd.save_hist) write_state((long)i, u, v, a);
            }
        }
        solved = true;
        dyn_solved = c.dynamic;
        nl_solved = c.nonlinear;
    }

    // Static solve with nonlinearGeom (Model::solve, S/model_analysis.rs:
    // 625-646): per staticLoadTime, loadRampSteps steps with factors
    // 1 - ((L - 1 - s) / L)^2, each a Newton loop (max|du| <= 1e-12, at most
    // 50 iterations) from the current state; brick8 and brickIM.  Each Newton
    // system is solved as the linear ones: the exact LDL^T factor of the new
    // tangent (direct) or conj_grad_sparse with its block factor (iterative).
    void solve_nonlinear() {
        const Command& c = solve_cmd;
        if (etype == ASEND_SHELL4) die("geometric nonlinearity is supported for brick8 and brickIM models");
        const bool iterative = (c.method == "iterative");
        check(asend_assemble(ctx, 0.0), "asend_assemble");
        asend_ldl_info li;
        std::memset(&li, 0, sizeof li);
        asend_solve_opts o;
        std::memset(&o, 0, sizeof o);
        o.check_every = 1;
        const int frc = asend_ldl_factor(ctx, iterative ? 6 * (int64_t)solver_block_dim : 0, &li);
        if (frc == ASEND_OK) {
            std::printf("maxBandwidth: %lld\n", (long long)li.max_bandwidth);
            o.method = iterative ? ASEND_SOLVE_PCG_FP64 : ASEND_SOLVE_DIRECT;
            o.precond = ASEND_PC_LDL;
        } else {
            if (frc != ASEND_ERR_UNSUPPORTED) check(frc, "asend_ldl_factor");
            std::printf("note: LDL^T factor not available on the device (%s); FP64 Jacobi PCG instead\n",
                        asend_last_error(ctx));
            o.method = ASEND_SOLVE_PCG_FP64;
            o.precond = ASEND_PC_JACOBI;
        }
        if (iterative) {
            o.stop_rule = ASEND_STOP_REF_ABS;
            o.tol = c.conv_tol;
            o.max_it = c.max_it > 0 ? c.max_it : n_dof;
        } else if (o.method != ASEND_SOLVE_DIRECT) {
            o.stop_rule = ASEND_STOP_REL_RESIDUAL;
            o.tol = 1e-12;
            o.max_it = 100 * n_dof + 1000;
            o.check_every = 16;
        }
        asend_nl_opts nl;
        std::memset(&nl, 0, sizeof nl);
        nl.load_steps = (int)std::max<long>(1, c.load_ramp);
        nl.max_newton = 50;
        nl.du_tol = 1e-12;
        for (size_t i = 0; i < c.load_times.size(); ++i) {
            loads_at(c.load_times[i], f);
            asend_nl_stats st;
            std::memset(&st, 0, sizeof st);
            check(asend_static_nonlinear(ctx, f.data(), u.data(), &nl, &o, &st), "asend_static_nonlinear");
            std::printf("Nonlinear iterations: %d over %d load steps, last max solution step: %s\n",
                        st.newton_iterations, st.load_steps, fmt_display(st.last_du).c_str());
            if (st.unconverged_steps)
                std::printf("Warning: %d load step(s) ended at 50 nonlinear iterations\n", st.unconverged_steps);
            if (iterative) std::printf("Total CG iterations: %lld\n", (long long)st.linear_iterations);
            if (c.save_hist) write_state((long)i, u, v, a);
        }
    }

    void solve_dynamic() {
        const Command& c = solve_cmd;
        if (c.lump_mass) die("lumpMass is not supported by this solver");
        if (!(c.time_step > 0.0)) die("dynamic solves need timeStep > 0");
        check(asend_set_rayleigh_damping(ctx, c.ray_cm, c.ray_ck), "asend_set_rayleigh_damping");
        material_damping();
        check(asend_newmark_setup(ctx, c.time_step, c.beta, c.gamma), "asend_newmark_setup");
        if (c.save_hist) write_state(0, u, v, a);
        std::vector<double> u1(n_dof), v1(n_dof), a1(n_dof), fs;
        const double one = 1.0;
        double t = 0.0;
        long step = 1;
        int64_t its = 0;
        while (t < c.sim_period) {
            loads_at(t, fs);
            asend_newmark_stats st;
            std::memset(&st, 0, sizeof st);
            check(asend_newmark_run(ctx, 1, fs.data(), &one, u.data(), v.data(), a.data(), 0, nullptr, nullptr,
                                    u1.data(), v1.data(), a1.data(), 1e-10, 200000, &st),
                  "asend_newmark_run");
            if (st.unconverged_steps) std::printf("Warning: time step %ld: solve did not converge\n", step);
            its += st.total_iterations;
            u.swap(u1);
            v.swap(v1);
            a.swap(a1);
            if (c.save_hist) {
                write_state(step, u, v, a);
                ++step;
            }
            t += c.time_step;
        }
        f = fs;
        std::printf("device Newmark: %lld solver iterations\n", (long long)its);
    }

    // material damping matrices of the elements' section materials
    // (get_solid_damp_dfd0, S/element_properties.rs:1537-1575); brick8 only
    void material_damping() {
        bool any = false;
        for (const auto& m : materials) any = any || m.damped;
        if (!any) return;
        if (etype != ASEND_BRICK8) die("material damping is supported for brick8 models");
        std::vector<double> dm(36 * materials.size());
        for (size_t k = 0; k < materials.size(); ++k)
            std::copy(materials[k].damping, materials[k].damping + 36, dm.begin() + 36 * k);
        std::vector<int32_t> idx(conn.size(), 0);
        for (const auto& sec : sections) {
            int32_t mi = -1;
            for (size_t k = 0; k < materials.size(); ++k)
                if (materials[k].name == sec.mat) mi = (int32_t)k;
            if (mi < 0) die("material %s is not defined", sec.mat.c_str());
            for (int32_t e : element_set(sec.el_set)) idx[e] = mi;
        }
        check(asend_set_material_damping(ctx, (int64_t)materials.size(), dm.data(), idx.data()),
              "asend_set_material_damping");
    }

    // reaction forces: the element residual of build_elastic_soln_load(false)
    // negated (S/model_output.rs:206-216): K u, plus M a + D v for a dynamic
    // solve (D = rayDampCM M + rayDampCK K), without the penalty
    std::vector<double> reaction() {
        std::vector<double> rf(n_dof, 0.0);
        if (!solve_cmd.elastic) return rf;  // no displacements: no element forces
        if (nl_solved) {  // the internal force at the displacements (condensed for brickIM)
            check(asend_nl_assemble(ctx, u.data(), rf.data()), "asend_nl_assemble");
            return rf;
        }
        const double sf = asend_penalty_scale(ctx);
        if (!dyn_solved) {
            check(asend_spmv(ctx, 64, u.data(), rf.data()), "asend_spmv");
            for (int64_t d : pen) rf[d] -= sf * u[d];
            if (fth.size() == (size_t)n_dof)  // the element residual B^T C (strain - alpha T)
                for (int64_t i = 0; i < n_dof; ++i) rf[i] -= fth[i];
            return rf;
        }
        for (const auto& m : materials)
            if (m.damped) die("reactionForce of dynamic solves with material damping is not supported");
        std::vector<double> x1(n_dof), x2(n_dof), y(n_dof), y2(n_dof);
        for (int64_t i = 0; i < n_dof; ++i) {
            x1[i] = u[i] + solve_cmd.ray_ck * v[i];
            x2[i] = a[i] + solve_cmd.ray_cm * v[i];
        }
        check(asend_assemble(ctx, 0.0), "asend_assemble");
        check(asend_spmv(ctx, 64, x1.data(), rf.data()), "asend_spmv");
        for (int64_t d : pen) rf[d] -= sf * x1[d];
        check(asend_spmv(ctx, 64, x2.data(), y.data()), "asend_spmv");
        check(asend_assemble(ctx, 1.0), "asend_assemble");
        check(asend_spmv(ctx, 64, x2.data(), y2.data()), "asend_spmv");
        for (int64_t i = 0; i < n_dof; ++i) rf[i] += y2[i] - y[i];
        return rf;
    }

    // Objective terms (get_obj_val, S/objective_meth.rs:592-943): categories
    // displacement (the nodes of nodeSet, component) and stress, strain,
    // strainEnergy (brick8, at each element's centre, S/objective_meth.rs:
    // 665-735); operators powerNorm coef (sum (q - tgt)^p)^(1/p) (:329-337),
    // volumeIntegral coef (sum q V - tgt)^p and volumeAverage
    // coef (sum q V / sum V - tgt)^p (:422-480); a displacement term weights
    // every node by 1.  The q of a term, its targets and weights V follow.
    struct TermData {
        std::vector<double> q, tgt, vol;
        std::vector<int64_t> idx;  // DOF (displacement) or element
    };
    std::vector<double> cfields;  // centre fields of the current u (13 per element)
    bool cfields_ok = false;

    bool element_term(const ObjTerm& t) const {
        return t.category == "stress" || t.category == "strain" || t.category == "strainEnergy";
    }

    TermData term_data(const ObjTerm& t) {
        TermData d;
        if (t.category == "displacement") {
            if (t.component < 1 || t.component > dpn) die("objective component out of range");
            for (int32_t n : node_set(t.nd_set, false)) {
                d.idx.push_back(dof_index[6 * n + t.component - 1]);
                d.q.push_back(u[d.idx.back()]);
                d.vol.push_back(1.0);
            }
        } else if (element_term(t)) {
            if (etype != ASEND_BRICK8) die("objective terms of category %s are implemented for brick8 models",
                                           t.category.c_str());
            if (t.category != "strainEnergy" && (t.component < 1 || t.component > 6))
                die("objective component out of range");
            if (!cfields_ok) {
                cfields.assign(13 * conn.size(), 0.0);
                check(asend_element_centre_fields(ctx, u.data(), cfields.data()), "asend_element_centre_fields");
                cfields_ok = true;
            }
            for (int32_t e : element_set(t.el_set)) {
                const double* f = &cfields[13 * (size_t)e];
                double q;
                if (t.category == "stress") q = f[6 + t.component - 1];
                else if (t.category == "strain") q = f[t.component - 1];
                else {
                    q = 0.0;
                    for (int k = 0; k < 6; ++k) q += f[6 + k] * f[k];
                    q *= 0.5;
                }
                d.idx.push_back(e);
                d.q.push_back(q);
                d.vol.push_back(f[12]);
            }
        } else {
            die("objective terms of category %s are not supported by this solver", t.category.c_str());
        }
        if (t.tgt.size() > 1 && t.tgt.size() != d.q.size()) die("target list does not match the set");
        d.tgt.assign(d.q.size(), t.tgt.empty() ? 0.0 : t.tgt[0]);
        if (t.tgt.size() > 1) d.tgt = t.tgt;
        return d;
    }

    // value of a term and dL/dq of each of its entries
    double term_value(const ObjTerm& t, const TermData& d, std::vector<double>* w) {
        const size_t n = d.q.size();
        if (w) w->assign(n, 0.0);
        if (t.optr == "powerNorm") {
            double ps = 0.0;
            for (size_t k = 0; k < n; ++k) ps += std::pow(d.q[k] - d.tgt[k], t.expnt);
            if (w) {
                const double sc = t.coef * std::pow(ps, 1.0 / t.expnt - 1.0);
                for (size_t k = 0; k < n; ++k) (*w)[k] = sc * std::pow(d.q[k] - d.tgt[k], t.expnt - 1.0);
            }
            return t.coef * std::pow(ps, 1.0 / t.expnt);
        }
        if (t.optr == "volumeIntegral" || t.optr == "volumeAverage") {
            double vi = 0.0, tv = 0.0;
            for (size_t k = 0; k < n; ++k) {
                vi += d.q[k] * d.vol[k];
                tv += d.vol[k];
            }
            const double tg = t.tgt.empty() ? 0.0 : t.tgt[0];
            const bool avg = t.optr == "volumeAverage";
            const double err = (avg ? vi / tv : vi) - tg;
            if (w) {
                const double sc = t.coef * t.expnt * std::pow(err, t.expnt - 1.0) / (avg ? tv : 1.0);
                for (size_t k = 0; k < n; ++k) (*w)[k] = sc * d.vol[k];
            }
            return t.coef * std::pow(err, t.expnt);
        }
        die("objective operator %s is not supported by this solver", t.optr.c_str());
    }

    void objective_values() {
        if (!solved) die("the objective needs a solution (run solve first)");
        if (dyn_solved) die("objectives of dynamic solves are not supported by this solver");
        if (nl_solved) die("objectives of nonlinear solves are not supported by this solver");
        if (!solve_cmd.elastic) die("objectives of thermal solves are not supported by this solver");
        for (const auto& m : materials)
            if (solve_cmd.thermal && m.expands)
                die("objectives of solves with thermal expansion are not supported by this solver");
        cfields_ok = false;
        for (auto& t : obj) {
            const TermData d = term_data(t);
            t.value = term_value(t, d, nullptr);
        }
    }

    // isotropic C of (E, nu) and dC/dp = -C (dS/dp) C for the nine compliance
    // parameters p = E1, E2, E3, nu12, nu13, nu23, G12, G13, G23 at that point
    // (the compliance of get_solid_stiff_dfd0, S/element_properties.rs:728-759)
    static void iso_c(double E, double nu, double* C) {
        const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
        for (int i = 0; i < 36; ++i) C[i] = 0.0;
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) C[6 * i + j] = lam;
            C[7 * i] += 2.0 * mu;
            C[7 * (3 + i)] = mu;
        }
    }
    static void dc_dp(double E, double nu, int p, double* dC) {
        const double G = E / (2.0 * (1.0 + nu));
        double dS[36] = {}, C[36], T[36];
        auto sym = [&](int i, int j, double v) { dS[6 * i + j] = dS[6 * j + i] = v; };
        switch (p) {
            case 0: sym(0, 0, -1.0 / (E * E)); sym(0, 1, nu / (E * E)); sym(0, 2, nu / (E * E)); break;
            case 1: sym(1, 1, -1.0 / (E * E)); sym(1, 2, nu / (E * E)); break;
            case 2: sym(2, 2, -1.0 / (E * E)); break;
            case 3: sym(0, 1, -1.0 / E); break;
            case 4: sym(0, 2, -1.0 / E); break;
            case 5: sym(1, 2, -1.0 / E); break;
            default: sym(p - 3, p - 3, -1.0 / (G * G)); break;
        }
        iso_c(E, nu, C);
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) {
                double a = 0.0;
                for (int k = 0; k < 6; ++k) a += dS[6 * i + k] * C[6 * k + j];
                T[6 * i + j] = a;
            }
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) {
                double a = 0.0;
                for (int k = 0; k < 6; ++k) a += C[6 * i + k] * T[6 * k + j];
                dC[6 * i + j] = -a;
            }
    }

    // adjoint gradient (get_obj_gradient, S/model_analysis.rs:1544-1706):
    // dL/du of the terms (getd_ld_u: displacement entries directly, the
    // element entries through B_c^T of their strain-space weights,
    // S/objective_meth.rs:944-1160), K lambda = dL/du, then per design
    // variable the element partials of -lambda^T (dK/dp) u times the DV
    // coefficients plus the explicit dL/dp of stress and strainEnergy terms
    // (getd_ld_d, S/objective_meth.rs:1372-1500: dsigma/dp = dC/dp eps)
    void gradient() {
        objective_values();
        shape_ok = false;
        std::vector<double> dLdu(n_dof, 0.0);
        std::vector<TermData> tds;
        std::vector<std::vector<double>> tws;
        std::vector<double> h;  // strain-space weights of the element terms
        for (const auto& t : obj) {
            tds.push_back(term_data(t));
            tws.emplace_back();
            term_value(t, tds.back(), &tws.back());
            const TermData& d = tds.back();
            const std::vector<double>& w = tws.back();
            if (t.category == "displacement") {
                for (size_t k = 0; k < d.q.size(); ++k) dLdu[d.idx[k]] += w[k];
                continue;
            }
            if (h.empty()) h.assign(6 * conn.size(), 0.0);
            for (size_t k = 0; k < d.q.size(); ++k) {
                const size_t e = (size_t)d.idx[k];
                const double* f = &cfields[13 * e];
                double* he = &h[6 * e];
                if (t.category == "strain") {
                    he[t.component - 1] += w[k];
                } else if (t.category == "stress") {
                    double C[36];
                    iso_c(E[e], nu[e], C);
                    for (int j = 0; j < 6; ++j) he[j] += w[k] * C[6 * (t.component - 1) + j];
                } else {
                    for (int j = 0; j < 6; ++j) he[j] += w[k] * f[6 + j];
                }
            }
        }
        if (!h.empty()) {
            std::vector<double> y(n_dof);
            check(asend_element_centre_adjoint(ctx, h.data(), y.data()), "asend_element_centre_adjoint");
            for (int64_t i = 0; i < n_dof; ++i) dLdu[i] += y[i];
        }
        dLdD.assign(dvars.size(), 0.0);
        if (!dvars.empty()) {
            if (etype != ASEND_BRICK8) die("design-variable gradients are implemented for brick8 models");
            std::vector<double> lam;
            device_solve(dLdu, lam, false);
            std::vector<double> part(9 * conn.size());
            check(asend_element_sens_material(ctx, u.data(), lam.data(), part.data()), "asend_element_sens_material");
            for (size_t i = 0; i < dvars.size(); ++i) {
                const auto& dv = dvars[i];
                if (dv.category == "density") {
                    dLdD[i] = density_gradient(dv, lam);
                    continue;
                }
                if (dv.category == "nodeCoord") {
                    dLdD[i] = coord_gradient(dv, lam);
                    continue;
                }
                const int base = dv.category == "modulus" ? 0 : dv.category == "poissonRatio" ? 3 : 6;
                const int p = base + dv.component - 1;
                const auto& set = element_set(dv.el_set);
                std::vector<double> cw(conn.size(), 0.0);  // DV coefficient of each element
                double acc = 0.0;
                for (size_t k = 0; k < set.size(); ++k) {
                    const double c = dv.coefs.empty() ? 1.0 : (dv.coefs.size() == 1 ? dv.coefs[0] : dv.coefs[k]);
                    acc += c * part[9 * (size_t)set[k] + p];
                    cw[set[k]] += c;
                }
                // explicit dependence of the element terms on the material
                for (size_t ti = 0; ti < obj.size(); ++ti) {
                    const auto& t = obj[ti];
                    if (t.category != "stress" && t.category != "strainEnergy") continue;
                    const TermData& d = tds[ti];
                    for (size_t k = 0; k < d.q.size(); ++k) {
                        const size_t e = (size_t)d.idx[k];
                        if (cw[e] == 0.0) continue;
                        const double* f = &cfields[13 * e];
                        double dC[36], ds[6];
                        dc_dp(E[e], nu[e], p, dC);
                        for (int r = 0; r < 6; ++r) {
                            ds[r] = 0.0;
                            for (int c = 0; c < 6; ++c) ds[r] += dC[6 * r + c] * f[c];
                        }
                        double dq;
                        if (t.category == "stress") {
                            dq = ds[t.component - 1];
                        } else {
                            dq = 0.0;
                            for (int r = 0; r < 6; ++r) dq += ds[r] * f[r];
                            dq *= 0.5;
                        }
                        acc += cw[e] * tws[ti][k] * dq;
                    }
                }
                dLdD[i] = acc;
            }
        }
        have_grad = true;
    }

    // dL/dD of a nodeCoord DV: -lam^T dR/dD with dR/dD = sum over the set's
    // nodes of coefficient x d(K u)/dX (the device's element shape
    // sensitivities; constraints do not depend on the coordinates).  Loads
    // that depend on the geometry (everything but nodalForce) are refused.
    std::vector<double> shape_s;
    bool shape_ok = false;
    double coord_gradient(const DesignVar& dv, const std::vector<double>& lam) {
        const double t = solve_cmd.load_times.empty() ? 0.0 : solve_cmd.load_times.back();
        for (const auto& l : loads)
            if (t >= l.active[0] && t <= l.active[1] && l.type != "nodalForce")
                die("nodeCoord design variables with %s loads are not supported", l.type.c_str());
        if (!shape_ok) {
            shape_s.assign(24 * conn.size(), 0.0);
            check(asend_element_sens_coords(ctx, u.data(), lam.data(), shape_s.data()), "asend_element_sens_coords");
            shape_ok = true;
        }
        std::vector<double> w(xyz.size() / 3, 0.0);  // coefficient of each node
        const auto& set = node_set(dv.nd_set, false);
        for (size_t k = 0; k < set.size(); ++k)
            w[set[k]] += dv.coefs.empty() ? 1.0 : (dv.coefs.size() == 1 ? dv.coefs[0] : dv.coefs[k]);
        double g = 0.0;
        for (size_t e = 0; e < conn.size(); ++e)
            for (int a = 0; a < 8; ++a) {
                const double c = w[conn[e][a]];
                if (c != 0.0) g -= c * shape_s[24 * e + 3 * a + dv.component - 1];
            }
        return g;
    }

    // dL/dD of a density DV in a static solve: the stiffness does not depend
    // on density, the loads of the gravitational and centrifugal kind do,
    // linearly in each element's density, so dR/dD = -sum over those loads of
    // the load of the DV's elements computed with density = coefficient, and
    // dL/dD = -lam^T dR/dD (the DV loop of get_obj_gradient,
    // S/model_analysis.rs:1679-1701, with d_relasticd_d's load term
    // :1379-1457)
    double density_gradient(const DesignVar& dv, const std::vector<double>& lam) {
        const auto& set = element_set(dv.el_set);
        std::vector<double> w(conn.size(), 0.0);
        for (size_t k = 0; k < set.size(); ++k)
            w[set[k]] += dv.coefs.empty() ? 1.0 : (dv.coefs.size() == 1 ? dv.coefs[0] : dv.coefs[k]);
        const size_t nn = xyz.size() / 3, ne = conn.size();
        std::vector<int32_t> cflat(8 * ne);
        for (size_t e = 0; e < ne; ++e)
            for (int a = 0; a < 8; ++a) cflat[8 * e + a] = conn[e][a];
        const double t = solve_cmd.load_times.empty() ? 0.0 : solve_cmd.load_times.back();
        double g = 0.0;
        std::vector<double> df;
        for (const auto& l : loads) {
            if (!(t >= l.active[0] && t <= l.active[1])) continue;
            if (l.type != "gravitational" && l.type != "centrifugal") continue;
            const auto& lset = element_set(l.el_set);
            df.assign(n_dof, 0.0);
            const int rc = asend_body_load(etype, l.type == "gravitational" ? 1 : 2, (int64_t)nn, xyz_dv.data(),
                                           (int64_t)ne, cflat.data(), dof_index.data(), (int64_t)lset.size(),
                                           lset.data(), w.data(), l.load, l.center, l.axis, l.ang_vel, n_dof,
                                           df.data());
            if (rc != ASEND_OK) die("%s load failed (%d)", l.type.c_str(), rc);
            for (int64_t i = 0; i < n_dof; ++i) g += lam[i] * df[i];
        }
        return g;
    }

    // ------------------------------------------------------------ writers
    // write_node_results (S/model_output.rs:84-262): with timeSteps, one file
    // per listed step, <name up to its first '.'>_timestep<i><rest>, from
    // that step's state file (which then stays the current state, as the
    // reference's backstep does), the time column being timeStep * i
    // (dynamic) or the i-th staticLoadTime; without, the current state and -1
    void write_node_results(const Command& c) {
        if (!solved) die("node results need a solution (run solve first)");
        for (const auto& fld : c.fields)
            if (fld != "displacement" && fld != "velocity" && fld != "acceleration" && fld != "reactionForce" &&
                fld != "temperature" && fld != "tdot" && fld != "reactionHeatGen")
                die("unrecognized output field '%s' for write_node_results", fld.c_str());
        if (c.time_steps.empty()) {
            write_node_file(c, c.file, -1.0);
            return;
        }
        const size_t dot = c.file.find('.');
        const std::string root = dot == std::string::npos ? c.file : c.file.substr(0, dot);
        const std::string ext = dot == std::string::npos ? "" : c.file.substr(dot);
        for (long ts : c.time_steps) {
            read_state(ts, u, v, a);
            double time;
            if (dyn_solved) {
                time = solve_cmd.time_step * (double)ts;
            } else {
                const auto& lt = solve_cmd.load_times;
                time = lt[std::min<size_t>((size_t)std::max<long>(ts, 0), lt.size() - 1)];
            }
            write_node_file(c, root + "_timestep" + std::to_string(ts) + ext, time);
        }
    }

    void write_node_file(const Command& c, const std::string& name, double time) {
        const auto& set = node_set(c.node_set, true);
        std::vector<double> rf, rh;
        for (const auto& fld : c.fields) {
            if (fld == "reactionForce" && rf.empty()) rf = reaction();
            if (fld == "reactionHeatGen" && rh.empty()) rh = reaction_heat();
        }
        if (temp.size() != xyz.size() / 3) temp.assign(xyz.size() / 3, 0.0);
        if (tdot.size() != xyz.size() / 3) tdot.assign(xyz.size() / 3, 0.0);
        std::FILE* fp = std::fopen(name.c_str(), "w");
        if (!fp) die("could not open file %s", name.c_str());
        std::fprintf(fp, "node,");
        for (const auto& fld : c.fields) {
            const char* h = fld == "displacement"      ? "U1,U2,U3,R1,R2,R3,"
                            : fld == "velocity"        ? "V1,V2,V3,RV1,RV2,RV3,"
                            : fld == "acceleration"    ? "A1,A2,A3,RA1,RA2,RA3,"
                            : fld == "temperature"     ? "T,"
                            : fld == "tdot"            ? "TDOT,"
                            : fld == "reactionHeatGen" ? "RHG,"
                                                       : "RF1,RF2,RF3,RM1,RM2,RM3,";
            std::fprintf(fp, "%s", h);
        }
        std::fprintf(fp, "time\n");
        for (int32_t n : set) {
            std::fprintf(fp, "%d,", n);
            for (const auto& fld : c.fields) {
                if (fld == "temperature" || fld == "tdot") {
                    std::fprintf(fp, "%s,", fmt_e12(fld == "temperature" ? temp[n] : tdot[n]).c_str());
                    continue;
                }
                if (fld == "reactionHeatGen") {
                    std::fprintf(fp, "%s,", fmt_e12(rh[dof_index[6 * n]]).c_str());
                    continue;
                }
                const std::vector<double>& vec = fld == "displacement"   ? u
                                                 : fld == "velocity"     ? v
                                                 : fld == "acceleration" ? a
                                                                         : rf;
                for (int k = 0; k < 6; ++k) {
                    const int64_t d = (k < dpn) ? dof_index[6 * n + k] : -1;
                    std::fprintf(fp, "%s,", fmt_e12(d >= 0 ? vec[d] : 0.0).c_str());
                }
            }
            std::fprintf(fp, "%s\n", fmt_e12(time).c_str());
        }
        std::fclose(fp);
    }

    // write_element_results (S/model_output.rs:264-470) at the element
    // centroids of solid elements: strain (engineering shears; Green-Lagrange
    // after a nonlinearGeom solve, get_solid_strain_dfd0,
    // S/element_soln_fields.rs:1345-1461) with its principal values, stress
    // C strain with von Mises and principal values (Model::mises,
    // principal_stress / principal_strain, S/model_res_utils.rs:139-196,
    // ascending), strainEnergyDen = stress . strain / 2.  The bubbles of
    // brickIM have zero gradient at the centroid.  Host arithmetic on the
    // solved displacements (output, not the solve path); the principal values
    // come from Jacobi rotations, the reference's from its own iteration.
    void write_element_results(const Command& c) {
        if (!solved) die("element results need a solution (run solve first)");
        if (etype == ASEND_SHELL4) die("element results of shell elements are not supported by this solver");
        if (c.position != "centroid") die("element results at position %s are not supported", c.position.c_str());
        for (const auto& m : materials)
            if (solve_cmd.thermal && m.expands)
                die("element results of solves with thermal expansion are not supported by this solver");
        for (const auto& sec : sections)
            if (sec.oriented)
                die("element results in oriented sections (local-frame output) are not supported by this solver");
        for (const auto& fld : c.fields)
            if (fld != "stress" && fld != "strain" && fld != "strainEnergyDen")
                die("element result field %s is not supported by this solver", fld.c_str());
        if (c.time_steps.empty()) {
            write_element_file(c, c.file, -1.0);
            return;
        }
        const size_t dot = c.file.find('.');
        const std::string root = dot == std::string::npos ? c.file : c.file.substr(0, dot);
        const std::string ext = dot == std::string::npos ? "" : c.file.substr(dot);
        for (long ts : c.time_steps) {
            read_state(ts, u, v, a);
            double time;
            if (dyn_solved) {
                time = solve_cmd.time_step * (double)ts;
            } else {
                const auto& lt = solve_cmd.load_times;
                time = lt[std::min<size_t>((size_t)std::max<long>(ts, 0), lt.size() - 1)];
            }
            write_element_file(c, root + "_timestep" + std::to_string(ts) + ext, time);
        }
    }

    // eigenvalues of a symmetric 3x3 matrix by cyclic Jacobi rotations, ascending
    static void sym_eig3(const double* m_in, double* ev) {
        double m[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) m[i][j] = m_in[3 * i + j];
        for (int sweep = 0; sweep < 50; ++sweep) {
            const double off = m[0][1] * m[0][1] + m[0][2] * m[0][2] + m[1][2] * m[1][2];
            const double dg = m[0][0] * m[0][0] + m[1][1] * m[1][1] + m[2][2] * m[2][2];
            if (off <= 1e-36 * dg || off == 0.0) break;
            for (int p = 0; p < 2; ++p)
                for (int q = p + 1; q < 3; ++q) {
                    if (m[p][q] == 0.0) continue;
                    const double th = (m[q][q] - m[p][p]) / (2.0 * m[p][q]);
                    const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                    const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
                    for (int k = 0; k < 3; ++k) {  // columns p, q
                        const double kp = m[k][p], kq = m[k][q];
                        m[k][p] = cs * kp - sn * kq;
                        m[k][q] = sn * kp + cs * kq;
                    }
                    for (int k = 0; k < 3; ++k) {  // rows p, q
                        const double pk = m[p][k], qk = m[q][k];
                        m[p][k] = cs * pk - sn * qk;
                        m[q][k] = sn * pk + cs * qk;
                    }
                }
        }
        for (int i = 0; i < 3; ++i) ev[i] = m[i][i];
        std::sort(ev, ev + 3);
    }

    void write_element_file(const Command& c, const std::string& name, double time) {
        const auto& set = element_set(c.el_set);
        std::FILE* fp = std::fopen(name.c_str(), "w");
        if (!fp) die("could not open file %s", name.c_str());
        std::fprintf(fp, "element,int_pt,layer,");
        for (const auto& fld : c.fields)
            std::fprintf(fp, "%s", fld == "stress"   ? "S11,S22,S33,S12,S13,S23,MISES,PS1,PS2,PS3,"
                                   : fld == "strain" ? "E11,E22,E33,E12,E13,E23,PE1,PE2,PE3,"
                                                     : "SE,");
        std::fprintf(fp, "time\n");
        static const double sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                        {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
        for (int32_t e : set) {
            double X[8][3], U[8][3];
            for (int t = 0; t < 8; ++t)
                for (int k = 0; k < 3; ++k) {
                    const int32_t nd = conn[e][t];
                    X[t][k] = xyz_dv[3 * (size_t)nd + k];
                    const int64_t d = dof_index[6 * (size_t)nd + k];
                    U[t][k] = d >= 0 ? u[d] : 0.0;
                }
            // centroid: dN_t/ds_j = sg[t][j] / 8
            double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            for (int t = 0; t < 8; ++t)
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) J[i][j] += X[t][i] * 0.125 * sg[t][j];
            const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2],
                         c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
            const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
            const double Ji[3][3] = {{c00 / det, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det,
                                      (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det},
                                     {c01 / det, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det,
                                      (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det},
                                     {c02 / det, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det,
                                      (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det}};
            double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // H_ij = du_i / dx_j
            for (int t = 0; t < 8; ++t) {
                double g[3];
                for (int j = 0; j < 3; ++j)
                    g[j] = 0.125 * (sg[t][0] * Ji[0][j] + sg[t][1] * Ji[1][j] + sg[t][2] * Ji[2][j]);
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) H[i][j] += U[t][i] * g[j];
            }
            const int pr[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
            double eps[6], sig[6];
            for (int q = 0; q < 6; ++q) {
                const int i = pr[q][0], j = pr[q][1];
                double v = H[i][j] + H[j][i];
                if (solve_cmd.nonlinear)
                    for (int k = 0; k < 3; ++k) v += H[k][i] * H[k][j];
                eps[q] = q < 3 ? 0.5 * v : v;
            }
            if (!cmats.empty()) {
                const double* C = cmats.data() + 36 * sec_of_el[e];
                for (int q = 0; q < 6; ++q) {
                    sig[q] = 0.0;
                    for (int r = 0; r < 6; ++r) sig[q] += E[e] * C[6 * q + r] * eps[r];
                }
            } else {
                const double Ee = E[e], ne = nu[e];
                const double lam = Ee * ne / ((1.0 + ne) * (1.0 - 2.0 * ne)), mu = Ee / (2.0 * (1.0 + ne));
                const double tr = eps[0] + eps[1] + eps[2];
                for (int q = 0; q < 6; ++q) sig[q] = q < 3 ? lam * tr + 2.0 * mu * eps[q] : mu * eps[q];
            }
            std::fprintf(fp, "%d,0,0,", e);
            for (const auto& fld : c.fields) {
                if (fld == "strain") {
                    for (int q = 0; q < 6; ++q) std::fprintf(fp, "%s,", fmt_e12(eps[q]).c_str());
                    const double m[9] = {eps[0],       0.5 * eps[3], 0.5 * eps[4], 0.5 * eps[3], eps[1],
                                         0.5 * eps[5], 0.5 * eps[4], 0.5 * eps[5], eps[2]};
                    double pv[3];
                    sym_eig3(m, pv);
                    for (int k = 0; k < 3; ++k) std::fprintf(fp, "%s,", fmt_e12(pv[k]).c_str());
                } else if (fld == "stress") {
                    for (int q = 0; q < 6; ++q) std::fprintf(fp, "%s,", fmt_e12(sig[q]).c_str());
                    const double a1 = sig[0] - sig[1], b1 = sig[1] - sig[2], c1 = sig[2] - sig[0];
                    const double d1 = sig[3] * sig[3] + sig[4] * sig[4] + sig[5] * sig[5];
                    std::fprintf(fp, "%s,", fmt_e12(std::sqrt(0.5 * (a1 * a1 + b1 * b1 + c1 * c1 + 6.0 * d1))).c_str());
                    const double m[9] = {sig[0], sig[3], sig[4], sig[3], sig[1], sig[5], sig[4], sig[5], sig[2]};
                    double pv[3];
                    sym_eig3(m, pv);
                    for (int k = 0; k < 3; ++k) std::fprintf(fp, "%s,", fmt_e12(pv[k]).c_str());
                } else {
                    double se = 0.0;
                    for (int q = 0; q < 6; ++q) se += sig[q] * eps[q];
                    std::fprintf(fp, "%s,", fmt_e12(0.5 * se).c_str());
                }
            }
            std::fprintf(fp, "%s\n", fmt_e12(time).c_str());
        }
        std::fclose(fp);
    }

    // Model::eigen_solve (S/model_analysis.rs:713-888) through asend_eigen:
    // the tangent stiffness at the current displacements (zero before any
    // solve), type frequency (lumped mass) or buckling (identity mass, load
    // factors), numModes, targetEigenvalue as the shift; the reference's
    // eigensolve runs for solverMethod direct only.
    std::vector<double> eig_vals, eig_fac, eig_vecs;
    Command modal_cmd;
    void modal_analysis(const Command& c) {
        const int nm = (int)c.num_modes;
        if (nm < 1) die("numModes must be positive");
        modal_cmd = c;
        if (c.modal_type != "buckling" && c.modal_type != "frequency") {
            // the reference solves only for these two types; any other leaves
            // the eigenvalues at zero (and the frequencies at sqrt(0) / 2 pi)
            std::printf("note: modal analysis type %s: no eigensolve\n", c.modal_type.c_str());
            eig_vals.assign(nm, 0.0);
            eig_fac.assign(nm, 0.0);
            eig_vecs.assign((size_t)nm * std::max<int64_t>(n_dof, 1), 0.0);
            return;
        }
        if (!c.method.empty() && c.method != "direct")
            die("modal analysis runs with solverMethod direct (the reference's other branch computes nothing)");
        if (!solved || !solve_cmd.elastic) {
            Command pc = c;
            pc.method = "direct";
            pc.elastic = true;
            pc.thermal = false;
            solve_cmd = pc;
            analysis_prep(pc);
            u.assign(n_dof, 0.0);
            v.assign(n_dof, 0.0);
            a.assign(n_dof, 0.0);
        }
        check(asend_assemble(ctx, 0.0), "asend_assemble");
        asend_ldl_info li;
        std::memset(&li, 0, sizeof li);
        check(asend_ldl_factor(ctx, 0, &li), "asend_ldl_factor");
        if (etype == ASEND_BRICKIM && !nl_solved) {
            // the elements hold the internal DOFs of the last linear solve
            std::vector<double> al(9 * conn.size());
            check(asend_brickim_internal(ctx, u.data(), al.data()), "asend_brickim_internal");
            check(asend_brickim_set_internal(ctx, al.data()), "asend_brickim_set_internal");
        }
        asend_solve_opts o;
        std::memset(&o, 0, sizeof o);
        o.method = ASEND_SOLVE_DIRECT;
        o.precond = ASEND_PC_LDL;
        eig_vals.assign(nm, 0.0);
        eig_fac.assign(nm, 0.0);
        eig_vecs.assign((size_t)nm * n_dof, 0.0);
        asend_modal_stats st;
        std::memset(&st, 0, sizeof st);
        check(asend_eigen(ctx, c.modal_type == "buckling" ? 1 : 0, u.data(), nm, c.tgt_eval, &o, eig_vals.data(),
                          eig_fac.data(), eig_vecs.data(), &st),
              "asend_eigen");
        std::printf("finished eigensolve: %lld Lanczos steps, %lld modes\n", (long long)st.lanczos_steps,
                    (long long)st.modes_found);
    }

    // Model::read_node_results (S/model_input.rs:1403-1466): a node results
    // CSV (header row with "node", then label,values...) sets the named
    // nodal fields (U1..R3, V1..RV3, A1..RA3, T, TDOT); other columns are
    // ignored.  The values apply now and to the next solve's starting state.
    std::vector<double> nd_state;  // per node: 6 u, 6 v, 6 a, T, Tdot

    // Model::read_initial_state (S/model_input.rs:1125-1200), applied by the
    // first analysis prep (Node::initialize_disp / initialize_temp,
    // S/node_meth.rs:104-119, S/model_prep.rs:564,598): "initialState" rows of
    // a node set and up to six values for displacement, velocity and
    // acceleration (missing ones zero), or one value for temperature and tdot.
    std::vector<double> init_state;  // per node, as nd_state
    bool init_pending = false, first_solve_done = false;
    void read_initial_state(const std::string& path) {
        LineReader lr;
        std::vector<std::string> d;
        const size_t nn = xyz.size() / 3;
        for (const auto& line : read_lines(path, "initial state")) {
            lr.read(line, d);
            const std::string* h = lr.head;
            if (h[0] != "initialState") continue;
            int off = -1;
            size_t nv = 0;
            if ((h[1] == "displacement" || h[1] == "velocity" || h[1] == "acceleration") && d.size() > 3) {
                off = h[1] == "displacement" ? 0 : h[1] == "velocity" ? 6 : 12;
                nv = 6;
            } else if ((h[1] == "temperature" || h[1] == "tdot") && d.size() == 2) {
                off = h[1] == "temperature" ? 18 : 19;
                nv = 1;
            }
            if (off < 0) continue;
            if (init_state.size() != 20 * nn) init_state.resize(20 * nn, 0.0);
            for (int32_t n : node_set(d[0], false))
                for (size_t k = 0; k < nv; ++k) init_state[20 * (size_t)n + off + k] = 1 + k < d.size() ? to_d(d[1 + k]) : 0.0;
            init_pending = true;
        }
    }
    void read_node_results(const std::string& path) {
        const size_t nn = xyz.size() / 3;
        if (nd_state.size() != 20 * nn) nd_state.assign(20 * nn, 0.0);
        std::vector<std::string> head;
        static const char* names[20] = {"U1", "U2", "U3", "R1", "R2", "R3", "V1", "V2", "V3", "RV1",
                                        "RV2", "RV3", "A1", "A2", "A3", "RA1", "RA2", "RA3", "T", "TDOT"};
        auto split = [](const std::string& ln) {
            std::vector<std::string> out;
            size_t a = 0;
            while (true) {
                const size_t b = ln.find(',', a);
                out.push_back(ln.substr(a, b == std::string::npos ? std::string::npos : b - a));
                if (b == std::string::npos) break;
                a = b + 1;
            }
            return out;
        };
        for (const auto& line : read_lines(path, "node results")) {
            if (line.find("node") != std::string::npos) {
                head = split(line);
                continue;
            }
            if (line.find(',') == std::string::npos) continue;
            const auto d = split(line);
            const long n = to_i(d[0]);
            if (n < 0 || (size_t)n >= nn) die("problem reading the node results file, %s", path.c_str());
            for (size_t i = 1; i < d.size() && i < head.size(); ++i)
                for (int k = 0; k < 20; ++k)
                    if (head[i] == names[k]) nd_state[20 * (size_t)n + k] = to_d(d[i]);
        }
        apply_node_state();
        state_kept = true;
    }

    void apply_node_state() {
        if (nd_state.empty() || n_dof == 0) return;
        const size_t nn = xyz.size() / 3;
        u.assign(n_dof, 0.0);
        v.assign(n_dof, 0.0);
        a.assign(n_dof, 0.0);
        for (size_t n = 0; n < nn; ++n)
            for (int k = 0; k < dpn; ++k) {
                const int64_t dd = dof_index[6 * n + k];
                if (dd < 0) continue;
                u[dd] = nd_state[20 * n + k];
                v[dd] = nd_state[20 * n + 6 + k];
                a[dd] = nd_state[20 * n + 12 + k];
            }
        if (temp.size() == nn)
            for (size_t n = 0; n < nn; ++n) {
                temp[n] = nd_state[20 * n + 18];
                if (tdot.size() == nn) tdot[n] = nd_state[20 * n + 19];
                const int64_t d = dof_index[6 * n];
                if (d >= 0 && (int64_t)t3.size() == n_dof) t3[d] = temp[n];
                if (d >= 0 && (int64_t)t3dot.size() == n_dof) t3dot[d] = tdot[n];
            }
    }

    // Model::set_soln_to_mode (S/model_analysis.rs:890-934): the displacement,
    // velocity or acceleration field set to a mode scaled to maxAmplitude
    // (largest entry); the next static solve starts from it
    bool state_kept = false;
    void set_soln_to_mode(const Command& c) {
        if (c.soln_field != "displacement" && c.soln_field != "velocity" && c.soln_field != "acceleration") {
            std::printf("Warning: %s is not a valid field to Set to a mode. Aborting setSolnToMode()\n",
                        c.soln_field.c_str());
            return;
        }
        if (eig_vals.empty() || c.mode < 0 || (size_t)c.mode >= eig_vals.size())
            die("setSolnToMode needs mode %ld of a modal analysis", c.mode);
        const double* ph = eig_vecs.data() + (size_t)c.mode * n_dof;
        double mx = 0.0;
        for (int64_t i = 0; i < n_dof; ++i) mx = std::max(mx, std::fabs(ph[i]));
        const double sc = c.max_amp / mx;
        std::vector<double>& dst = c.soln_field == "displacement" ? u : c.soln_field == "velocity" ? v : a;
        dst.assign(n_dof, 0.0);
        for (int64_t i = 0; i < n_dof; ++i) dst[i] = sc * ph[i];
        state_kept = true;
    }

    // Model::zero_solution (S/model_analysis.rs:675-712): the listed fields
    // (and with displacement the elements' internal DOFs) set to zero
    void zero_solution(const Command& c) {
        for (const auto& fld : c.fields) {
            std::vector<double>* vec = fld == "displacement"   ? &u
                                       : fld == "velocity"     ? &v
                                       : fld == "acceleration" ? &a
                                                               : nullptr;
            if (vec) vec->assign(vec->size(), 0.0);
            if (fld == "temperature") {
                std::fill(temp.begin(), temp.end(), 0.0);
                std::fill(t3.begin(), t3.end(), 0.0);
            }
            if (fld == "tdot") {
                std::fill(tdot.begin(), tdot.end(), 0.0);
                std::fill(t3dot.begin(), t3dot.end(), 0.0);
            }
            if (fld == "displacement" && ctx && etype == ASEND_BRICKIM) {
                std::vector<double> zero(9 * conn.size(), 0.0);
                check(asend_brickim_set_internal(ctx, zero.data()), "asend_brickim_set_internal");
            }
        }
        state_kept = true;
    }

    // write_modal_results (S/model_output.rs:495-551): mode, eigenvalue and
    // load factor or frequency; with writeModes (the flag of the
    // modalAnalysis command, as the reference reads it) one file per mode,
    // <name up to its first '.'>_mode<i><rest>, Node,U1..R3 per node label
    // (solid nodes carry no rotations: 0 here, where the reference would
    // index its mode vector with the unset rotation DOF numbers)
    void write_modal_results(const Command& c) {
        if (eig_vals.empty()) die("modal results need a modal analysis (run modalAnalysis first)");
        std::FILE* fp = std::fopen(c.file.c_str(), "w");
        if (!fp) die("could not open file %s", c.file.c_str());
        const bool buck = modal_cmd.modal_type == "buckling";
        std::fprintf(fp, "mode,eigenvalue,%s\n", buck ? "load_factor" : "frequency");
        for (size_t i = 0; i < eig_vals.size(); ++i)
            std::fprintf(fp, "%zu,%s,%s\n", i, fmt_e12(eig_vals[i]).c_str(), fmt_e12(eig_fac[i]).c_str());
        std::fclose(fp);
        if (!modal_cmd.write_modes) return;
        const size_t dot = c.file.find('.');
        const std::string root = dot == std::string::npos ? c.file : c.file.substr(0, dot);
        const std::string ext = dot == std::string::npos ? "" : c.file.substr(dot);
        const size_t nn = xyz.size() / 3;
        for (size_t i = 0; i < eig_vals.size() && n_dof > 0; ++i) {
            const std::string name = root + "_mode" + std::to_string(i) + ext;
            std::FILE* mp = std::fopen(name.c_str(), "w");
            if (!mp) die("could not open file %s", name.c_str());
            std::fprintf(mp, "Node,U1,U2,U3,R1,R2,R3\n");
            for (size_t n = 0; n < nn; ++n) {
                std::fprintf(mp, "%zu", n);
                for (int k = 0; k < 6; ++k) {
                    const int64_t d = k < dpn ? dof_index[6 * n + k] : -1;
                    std::fprintf(mp, ",%s", fmt_e12(d >= 0 ? eig_vecs[i * (size_t)n_dof + d] : 0.0).c_str());
                }
                std::fprintf(mp, "\n");
            }
            std::fclose(mp);
        }
    }

    void write_objective(const Command& c) {
        std::FILE* fp = std::fopen(c.file.c_str(), "w");
        if (!fp) die("could not open file %s", c.file.c_str());
        std::fprintf(fp, "objectiveTerm,value");
        for (const auto& s : c.include) std::fprintf(fp, ",%s", s.c_str());
        std::fprintf(fp, "\n");
        for (size_t i = 0; i < obj.size(); ++i) {
            const auto& t = obj[i];
            std::fprintf(fp, "%zu,%s", i, fmt_e12(t.value).c_str());
            for (const auto& s : c.include) {
                if (s == "category") std::fprintf(fp, ",%s", t.category.c_str());
                else if (s == "operator") std::fprintf(fp, ",%s", t.optr.c_str());
                else if (s == "component") std::fprintf(fp, ",%d", t.component);
                else if (s == "coefficient") std::fprintf(fp, ",%s", fmt_e12(t.coef).c_str());
                else if (s == "exponent") std::fprintf(fp, ",%s", fmt_e12(t.expnt).c_str());
                else if (s == "elementSet") std::fprintf(fp, ",%s", t.el_set.c_str());
                else if (s == "nodeSet") std::fprintf(fp, ",%s", t.nd_set.c_str());
                else die("unrecognized field '%s' in the list for write_objective()", s.c_str());
            }
            std::fprintf(fp, "\n");
        }
        std::fclose(fp);
        if (c.write_grad) {
            const size_t dot = c.file.find('.');
            const std::string root = dot == std::string::npos ? c.file : c.file.substr(0, dot);
            const std::string ext = dot == std::string::npos ? "" : c.file.substr(dot);
            const std::string gname = root + "_grad" + ext;
            std::FILE* gp = std::fopen(gname.c_str(), "w");
            if (!gp) die("could not open file, %s", gname.c_str());
            std::fprintf(gp, "designVariable,objGrad\n");
            for (size_t i = 0; i < dvars.size(); ++i)
                std::fprintf(gp, "%zu,%s\n", i, fmt_e12(have_grad ? dLdD[i] : 0.0).c_str());
            std::fclose(gp);
        }
    }

    // ------------------------------------------------------------ driver
    void execute() {
        for (const auto& c : cmds) {
            if (c.cmd == "readModelInput") {
                std::printf("reading Model input: %s\n", c.file.c_str());
                read_model(c.file);
                read_constraints(c.file);
                read_loads(c.file);
                read_initial_state(c.file);
            } else if (c.cmd == "readInitialState") {
                std::printf("reading initial state: %s\n", c.file.c_str());
                read_initial_state(c.file);
            } else if (c.cmd == "readConstraints") {
                std::printf("reading constraints: %s\n", c.file.c_str());
                read_constraints(c.file);
            } else if (c.cmd == "readLoads") {
                std::printf("reading loads: %s\n", c.file.c_str());
                read_loads(c.file);
            } else if (c.cmd == "readDesignVarInput") {
                std::printf("reading design variable input: %s\n", c.file.c_str());
                read_dvars(c.file);
            } else if (c.cmd == "readDesignVarValues") {
                read_dvar_values(c.file);
            } else if (c.cmd == "readObjectiveInput") {
                std::printf("reading Objective input: %s\n", c.file.c_str());
                read_objective(c.file);
            } else if (c.cmd == "solvePrep") {
                // the ordering and upload happen with the solve
            } else if (c.cmd == "solve") {
                std::printf("running main analysis \n");
                solve(c);
            } else if (c.cmd == "calcObjective") {
                std::printf("calculating Objective function %s\n", c.file.c_str());
                objective_values();
            } else if (c.cmd == "calcObjGradient") {
                std::printf("calculating Objective gradient %s\n", c.file.c_str());
                gradient();
            } else if (c.cmd == "writeNodeResults") {
                std::printf("writing Node results\n");
                write_node_results(c);
            } else if (c.cmd == "readNodeResults") {
                std::printf("reading node results: %s\n", c.file.c_str());
                read_node_results(c.file);
            } else if (c.cmd == "setSolnToMode") {
                set_soln_to_mode(c);
            } else if (c.cmd == "zeroSolution") {
                zero_solution(c);
            } else if (c.cmd == "modalAnalysis") {
                std::printf("running modal analysis\n");
                modal_analysis(c);
            } else if (c.cmd == "writeModalResults") {
                std::printf("writing modal results\n");
                write_modal_results(c);
            } else if (c.cmd == "writeElementResults") {
                std::printf("writing Element results\n");
                write_element_results(c);
            } else if (c.cmd == "writeObjective") {
                std::printf("writing Objective results\n");
                write_objective(c);
            } else {
                die("job command %s is not supported by this solver", c.cmd.c_str());
            }
            std::fflush(stdout);
        }
    }
};

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) die("job input file not specified in call to analysis solver.");
    const bool dry = (argc > 2 && std::strcmp(argv[2], "--parse-only") == 0);
    std::printf("Beginning Job: %s\n", argv[1]);
    Job job;
    job.read_job(argv[1]);
    if (dry) {
        // parse the model files and report what was read, without a device
        for (const auto& c : job.cmds) {
            if (c.cmd == "readModelInput") {
                job.read_model(c.file);
                job.read_constraints(c.file);
                job.read_loads(c.file);
            } else if (c.cmd == "readConstraints") job.read_constraints(c.file);
            else if (c.cmd == "readLoads") job.read_loads(c.file);
            else if (c.cmd == "readDesignVarInput") job.read_dvars(c.file);
            else if (c.cmd == "readDesignVarValues") job.read_dvar_values(c.file);
            else if (c.cmd == "readObjectiveInput") job.read_objective(c.file);
        }
        std::printf("commands %zu nodes %zu elements %zu node_sets %zu element_sets %zu sections %zu materials %zu "
                    "constraints %zu loads %zu design_variables %zu objective_terms %zu temperature_constraints %zu "
                    "thermal_loads %zu\n",
                    job.cmds.size(), job.xyz.size() / 3, job.conn.size(), job.nsets.size(), job.esets.size(),
                    job.sections.size(), job.materials.size(), job.constraints.size(), job.loads.size(),
                    job.dvars.size(), job.obj.size(), job.tconstraints.size(), job.tloads.size());
        for (const auto& c : job.cmds) std::printf("command %s %s\n", c.cmd.c_str(), c.file.c_str());
        return 0;
    }
    job.execute();
    return 0;
}
