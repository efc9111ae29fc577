// BoltzTechs XML case layer: a small XML reader for the reference schema and the derivation
// of lattice-unit parameters (MaterKit/BoltzTechs*.xml; the reference reads them through
// XMLreader / readParameters, MaterKit/xml/mkTINYXMLIO.h:32-111, MaterKit/inc/mkFunction.h:62).
//
// Derived-parameter rules (the reference marks derived entries with negative values,
// e.g. MaterKit/BoltzTechs - thermal.xml:25-28,54):
//   lambda   = 0.8839 * W0 / d0                        (comment at thermal.xml:61)
//   alpha    = 0.6267 * lambda * W0^2 / t0 when alphaFluid <= 0
//   D~       = D * dt / dx^2          (lattice diffusivity per step, dx and dt in W0, t0)
//   omega    = 1 / (D~ / cs^2 + 1/2)  cs^2 = 1/4 on D3Q7 scalars, 1/3 on D3Q19/D3Q27
//   s_nu     = omega_f if omega_f > 0, else 1 / (3 nu dt/dx^2 + 1/2)
//   TRT      s_odd = 1 / (Lambda / (1/s_nu - 1/2) + 1/2)
// The optional <solver> block carries what the 2D reference schema has no tag for: lattices,
// operator, precision, per-face boundary codes, Rayleigh-Benard scaling and random seeds.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mkg.h"

namespace mkg {
void set_error(const std::string& msg);
}

namespace {

struct CaseErr {
    CaseErr& operator=(const std::string& m) {
        mkg::set_error(m);
        return *this;
    }
} g_case_err;

struct Node {
    std::string name, text;
    std::vector<std::unique_ptr<Node>> kids;
    const Node* child(const std::string& n) const {
        for (auto& k : kids)
            if (k->name == n) return k.get();
        return nullptr;
    }
};

// Minimal parser: elements, text, comments, declarations; no attributes are needed by the
// schema (attributes are skipped).
bool parse_xml(const std::string& s, Node& root, std::string& err) {
    std::vector<Node*> stack{&root};
    size_t i = 0, n = s.size();
    while (i < n) {
        if (s[i] != '<') {
            size_t j = s.find('<', i);
            if (j == std::string::npos) j = n;
            stack.back()->text += s.substr(i, j - i);
            i = j;
            continue;
        }
        if (s.compare(i, 4, "<!--") == 0) {
            size_t j = s.find("-->", i + 4);
            if (j == std::string::npos) { err = "unterminated comment"; return false; }
            i = j + 3;
            continue;
        }
        if (s.compare(i, 2, "<?") == 0 || s.compare(i, 2, "<!") == 0) {
            size_t j = s.find('>', i);
            if (j == std::string::npos) { err = "unterminated declaration"; return false; }
            i = j + 1;
            continue;
        }
        size_t j = s.find('>', i);
        if (j == std::string::npos) { err = "unterminated tag"; return false; }
        std::string tag = s.substr(i + 1, j - i - 1);
        i = j + 1;
        if (!tag.empty() && tag[0] == '/') {
            std::string name = tag.substr(1);
            while (!name.empty() && std::isspace((unsigned char)name.back())) name.pop_back();
            if (stack.size() < 2 || stack.back()->name != name) {
                err = "mismatched closing tag </" + name + ">";
                return false;
            }
            stack.pop_back();
            continue;
        }
        bool selfclose = !tag.empty() && tag.back() == '/';
        if (selfclose) tag.pop_back();
        size_t sp = 0;
        while (sp < tag.size() && !std::isspace((unsigned char)tag[sp])) sp++;
        auto node = std::make_unique<Node>();
        node->name = tag.substr(0, sp);
        Node* raw = node.get();
        stack.back()->kids.push_back(std::move(node));
        if (!selfclose) stack.push_back(raw);
    }
    if (stack.size() != 1) { err = "unclosed element <" + stack.back()->name + ">"; return false; }
    return true;
}

struct Case {
    Node root;
    const Node* top = nullptr;
    const Node* find(const char* path) const {
        const Node* n = top;
        std::string p(path);
        size_t a = 0;
        while (n && a <= p.size()) {
            size_t b = p.find('/', a);
            std::string seg = p.substr(a, b == std::string::npos ? std::string::npos : b - a);
            n = n->child(seg);
            if (b == std::string::npos) break;
            a = b + 1;
        }
        return n;
    }
    std::vector<std::string> tokens(const char* path) const {
        std::vector<std::string> out;
        const Node* n = find(path);
        if (!n) return out;
        std::istringstream is(n->text);
        std::string t;
        while (is >> t) out.push_back(t);
        return out;
    }
    bool has(const char* path) const { return find(path) != nullptr; }
    double num(const char* path, double dflt) const {
        auto t = tokens(path);
        if (t.empty()) return dflt;
        return std::strtod(t[0].c_str(), nullptr);
    }
    std::vector<double> nums(const char* path) const {
        std::vector<double> v;
        for (auto& t : tokens(path)) v.push_back(std::strtod(t.c_str(), nullptr));
        return v;
    }
    // case-insensitive boolean, as XMLreaderProxy::read<bool> (mkTINYXMLIO.hh:34-155)
    bool flag(const char* path, bool dflt) const {
        auto t = tokens(path);
        if (t.empty()) return dflt;
        std::string v = t[0];
        for (auto& ch : v) ch = (char)std::tolower((unsigned char)ch);
        return v == "true" || v == "1" || v == "yes";
    }
    std::string word(const char* path, const char* dflt) const {
        auto t = tokens(path);
        if (t.empty()) return dflt;
        std::string v = t[0];
        for (auto& ch : v) ch = (char)std::tolower((unsigned char)ch);
        return v;
    }
};

bool load_case(const char* path, Case& c) {
    std::ifstream f(path, std::ios::binary);
    if (!f) { g_case_err = std::string("cannot open case file ") + path; return false; }
    std::stringstream ss;
    ss << f.rdbuf();
    std::string err;
    if (!parse_xml(ss.str(), c.root, err)) { g_case_err = std::string(path) + ": " + err; return false; }
    c.top = c.root.child("MaterKit");
    if (!c.top) { g_case_err = std::string(path) + ": no <MaterKit> root element"; return false; }
    return true;
}

int bc_code(const std::string& w) {
    if (w == "periodic") return MKG_BC_PERIODIC;
    if (w == "bounce" || w == "wall" || w == "noflux") return MKG_BC_BOUNCE;
    if (w == "velocity") return MKG_BC_VELOCITY;
    if (w == "dirichlet") return MKG_BC_DIRICHLET;
    if (w == "outflow") return MKG_BC_OUTFLOW;
    if (w == "pressure") return MKG_BC_PRESSURE;
    return -1;
}

bool read_bc(const Case& c, const char* path, int32_t* bc) {
    auto t = c.tokens(path);
    if (t.empty()) return true;
    if (t.size() != 6) { g_case_err = std::string(path) + " needs 6 face codes"; return false; }
    for (int k = 0; k < 6; k++) {
        std::string w = t[k];
        for (auto& ch : w) ch = (char)std::tolower((unsigned char)ch);
        int v = bc_code(w);
        if (v < 0) { g_case_err = std::string(path) + ": unknown face code " + t[k]; return false; }
        bc[k] = v;
    }
    return true;
}

int lattice_q(const std::string& w) {
    if (w == "none" || w == "off") return 0;
    if (w == "d3q7") return 7;
    if (w == "d3q19") return 19;
    if (w == "d3q27") return 27;
    return -1;
}

uint64_t splitmix(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double uniform(uint64_t* s) { return (double)(splitmix(s) >> 11) * 0x1.0p-53; }

void euler_zyx_to_rot(double az, double ay, double ax, double* R) {
    // R = Rz(az) Ry(ay) Rx(ax), angles in degrees, anticlockwise positive
    const double d = M_PI / 180.0;
    double cz = std::cos(az * d), sz = std::sin(az * d), cy = std::cos(ay * d),
           sy = std::sin(ay * d), cx = std::cos(ax * d), sx = std::sin(ax * d);
    R[0] = cz * cy; R[1] = cz * sy * sx - sz * cx; R[2] = cz * sy * cx + sz * sx;
    R[3] = sz * cy; R[4] = sz * sy * sx + cz * cx; R[5] = sz * sy * cx - cz * sx;
    R[6] = -sy;     R[7] = cy * sx;                R[8] = cy * cx;
}

}  // namespace

extern "C" {

int mkg_quat_to_rot(const double* q, double* R) {
    if (!q || !R) return -1;
    double w = q[0], x = q[1], y = q[2], z = q[3];
    double n = std::sqrt(w * w + x * x + y * y + z * z);
    if (!(n > 0)) return -1;
    w /= n; x /= n; y /= n; z /= n;
    R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - w * z);     R[2] = 2 * (x * z + w * y);
    R[3] = 2 * (x * y + w * z);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - w * x);
    R[6] = 2 * (x * z - w * y);     R[7] = 2 * (y * z + w * x);     R[8] = 1 - 2 * (x * x + y * y);
    return 0;
}

// Centres uniform in [r, dim-1-r] per axis, then a unit quaternion per seed from three
// uniforms (Shoemake 1992): q = (sqrt(1-u1) sin 2pi u2, sqrt(1-u1) cos 2pi u2,
// sqrt(u1) sin 2pi u3, sqrt(u1) cos 2pi u3), stored as (w, x, y, z) = (q3, q0, q1, q2).
int mkg_random_seeds(uint64_t seed, int n, double radius, int nx, int ny, int nz,
                     double* centers, double* quats) {
    if (n < 0 || !centers) return -1;
    uint64_t s = seed;
    int dims[3] = {nx, ny, nz};
    for (int k = 0; k < n; k++)
        for (int a = 0; a < 3; a++) {
            double lo = radius, hi = dims[a] - 1 - radius;
            if (hi < lo) { lo = 0; hi = dims[a] - 1; }
            centers[3 * k + a] = lo + (hi - lo) * uniform(&s);
        }
    if (quats)
        for (int k = 0; k < n; k++) {
            double u1 = uniform(&s), u2 = uniform(&s), u3 = uniform(&s);
            double a = std::sqrt(1 - u1), b = std::sqrt(u1);
            double q0 = a * std::sin(2 * M_PI * u2), q1 = a * std::cos(2 * M_PI * u2);
            double q2 = b * std::sin(2 * M_PI * u3), q3 = b * std::cos(2 * M_PI * u3);
            quats[4 * k + 0] = q3;
            quats[4 * k + 1] = q0;
            quats[4 * k + 2] = q1;
            quats[4 * k + 3] = q2;
        }
    return 0;
}

int mkg_params_from_xml(const char* path, mkg_params* p, int32_t* run_info) {
    if (!path || !p) { g_case_err = "null argument"; return -1; }
    Case c;
    if (!load_case(path, c)) return -1;
    mkg_params_default(p);
    auto mesh = c.nums("geometry/MeshXYZ");
    if (mesh.size() != 3) { g_case_err = "geometry/MeshXYZ needs 3 values"; return -1; }
    p->nx = (int32_t)mesh[0];
    p->ny = (int32_t)mesh[1];
    p->nz = (int32_t)mesh[2];
    const double dx = c.num("numerics/mk_dx", 1.0), dt = c.num("numerics/mk_dt", 1.0);
    const double lu = dt / (dx * dx);

    const bool dendrite = c.flag("calcMode/dendrite", false);
    const bool flowf = c.flag("calcMode/flowField", false);
    p->flow_q = flowf ? 19 : 0;
    p->phase_q = dendrite ? 19 : 0;
    p->thermal = dendrite ? 1 : 0;
    p->solute = 0;
    if (c.has("solver/flowLattice")) p->flow_q = lattice_q(c.word("solver/flowLattice", "none"));
    if (c.has("solver/phaseLattice")) p->phase_q = lattice_q(c.word("solver/phaseLattice", "none"));
    if (c.has("solver/thermal")) p->thermal = c.flag("solver/thermal", false);
    if (c.has("solver/solute")) p->solute = c.flag("solver/solute", false);
    if (p->flow_q < 0 || p->phase_q < 0 || p->phase_q == 27) {
        g_case_err = "solver lattice must be none, D3Q7 (phase only), D3Q19 or D3Q27 (flow only)";
        return -1;
    }
    std::string prec = c.word("solver/precision", "fp64");
    p->precision = (prec == "fp32") ? MKG_PREC_FP32 : MKG_PREC_FP64;

    // ---- flow
    double sf = c.num("numerics/omega_f", -1.0);
    if (c.has("solver/tau")) sf = 1.0 / c.num("solver/tau", 1.0);
    const double nu_phys = c.num("phases/fluid/kinematicViscosity", 0.0);
    if (!(sf > 0) && nu_phys > 0) sf = 1.0 / (3.0 * nu_phys * lu + 0.5);
    if (!(sf > 0)) sf = 1.0;
    std::string op = c.word("solver/operator", "bgk");
    // equilibrium: compressible (default) or the He-Luo incompressible form
    const std::string eq = c.word("solver/equilibrium", "compressible");
    if (eq != "compressible" && eq != "incompressible") {
        g_case_err = "solver/equilibrium must be compressible or incompressible";
        return -1;
    }
    p->flow_incomp = eq == "incompressible" ? 1 : 0;
    // phase-field model: Karma-Rappel dendrite growth or the conservative Allen-Cahn interface
    const std::string pm = c.word("solver/phaseModel", "karmarappel");
    if (pm == "karmarappel") p->pf_model = MKG_PF_KARMA_RAPPEL;
    else if (pm == "allencahn") p->pf_model = MKG_PF_ALLEN_CAHN;
    else {
        g_case_err = "solver/phaseModel must be karmaRappel or allenCahn";
        return -1;
    }
    // anisotropy form of the phase field and its two parameters
    const std::string af = c.word("solver/anisotropyForm", "cubic");
    if (af == "cubic") p->aniso_kind = MKG_ANISO_CUBIC;
    else if (af == "cbc") p->aniso_kind = MKG_ANISO_CBC;
    else if (af == "hcp") p->aniso_kind = MKG_ANISO_HCP;
    else if (af == "facet") p->aniso_kind = MKG_ANISO_FACET;
    else {
        g_case_err = "solver/anisotropyForm must be cubic, cbc, hcp or facet";
        return -1;
    }
    auto ap = c.nums("solver/anisotropyParams");
    for (size_t k = 0; k < ap.size() && k < 2; k++) p->aniso[k] = ap[k];
    p->s_nu = sf;
    p->s_bulk = p->s_odd = p->s_even = sf;
    p->flow_op = MKG_OP_BGK;
    if (op == "trt") {
        const double lam = c.num("solver/magicLambda", 3.0 / 16.0);
        p->flow_op = MKG_OP_TRT;
        p->s_odd = 1.0 / (lam / (1.0 / sf - 0.5) + 0.5);
    } else if (op == "mrt") {
        p->flow_op = MKG_OP_MRT;
        auto r = c.nums("solver/mrtRates");
        p->s_bulk = r.size() > 0 ? r[0] : 1.19;
        p->s_odd = r.size() > 1 ? r[1] : 1.2;
        p->s_even = r.size() > 2 ? r[2] : 1.4;
    }
    auto acc = c.nums("flow/acceleration");
    if (acc.size() == 3)
        for (int k = 0; k < 3; k++) p->force[k] = acc[k];
    auto bf = c.nums("solver/bodyForce");
    if (bf.size() == 3)
        for (int k = 0; k < 3; k++) p->force[k] = bf[k];

    // ---- boundary codes
    if (!read_bc(c, "solver/bcFlow", p->bc_f) || !read_bc(c, "solver/bcPhase", p->bc_g) ||
        !read_bc(c, "solver/bcThermal", p->bc_h) || !read_bc(c, "solver/bcSolute", p->bc_c))
        return -1;

    // ---- phase field (Karma-Rappel units)
    const double W0 = c.num("phasechange/W0", 1.0), t0 = c.num("phasechange/t0", 1.0);
    const double d0 = c.num("phasechange/d0", 0.0);
    double lambda = c.num("phasechange/lambda", -1.0);
    if (!(lambda > 0) && d0 > 0) lambda = 0.8839 * W0 / d0;
    if (!(lambda > 0)) lambda = 0.0;
    p->lambda = lambda;
    p->eps4 = c.num("phasechange/anisotropy", 0.0);
    p->pf_dt = dt / t0;
    p->pf_wdx2 = (W0 / dx) * (W0 / dx);
    p->pf_coupling = p->phase_q ? (p->thermal ? 1 : (p->solute ? 2 : 0)) : 0;
    if (c.has("solver/coupling")) {
        std::string cp = c.word("solver/coupling", "none");
        p->pf_coupling = cp == "thermal" ? 1 : (cp == "solute" ? 2 : 0);
    }
    p->phi_s = c.num("solver/solidThreshold", 0.99);

    // ---- thermal
    double aF = c.num("temperature/alphaFluid", -1.0), aS = c.num("temperature/alphaSolid", -1.0);
    if (!(aF > 0)) aF = 0.6267 * lambda * W0 * W0 / t0;
    if (!(aS >= 0)) aS = aF;
    p->alpha_l = aF * lu;
    p->alpha_s = aS * lu;
    p->T_ref = c.num("temperature/iniValue", 0.0);
    p->ref_h = p->T_ref;
    p->latent_T = c.num("solver/latentHeat", 0.5);

    // Rayleigh-Benard scaling (z vertical): nu = u_ff H sqrt(Pr/Ra), alpha = nu/Pr,
    // g beta = u_ff^2 / (H dT), hot wall z-, cold wall z+.
    if (c.has("solver/rayleighBenard")) {
        const double Ra = c.num("solver/rayleighBenard/rayleigh", 1e5);
        const double Pr = c.num("solver/rayleighBenard/prandtl", 0.71);
        const double uff = c.num("solver/rayleighBenard/freeFall", 0.05);
        const double Th = c.num("solver/rayleighBenard/hot", 1.0);
        const double Tc = c.num("solver/rayleighBenard/cold", 0.0);
        const double H = p->nz;
        const double nu = uff * H * std::sqrt(Pr / Ra);
        const double al = nu / Pr;
        const double sn = 1.0 / (3.0 * nu + 0.5);
        const double lam = c.num("solver/magicLambda", 3.0 / 16.0);
        p->s_nu = p->s_bulk = p->s_even = sn;
        p->s_odd = (p->flow_op == MKG_OP_TRT) ? 1.0 / (lam / (1.0 / sn - 0.5) + 0.5) : sn;
        p->alpha_l = p->alpha_s = al;
        p->buoy[0] = p->buoy[1] = 0.0;
        p->buoy[2] = uff * uff / (H * (Th - Tc));
        p->T_ref = 0.5 * (Th + Tc);
        p->ref_h = p->T_ref;
        p->T_face[4] = Th;
        p->T_face[5] = Tc;
    }
    auto by = c.nums("solver/buoyancy");
    if (by.size() == 3)
        for (int k = 0; k < 3; k++) p->buoy[k] = by[k];
    auto tf = c.nums("solver/faceTemperature");
    if (tf.size() == 6)
        for (int k = 0; k < 6; k++) p->T_face[k] = tf[k];

    // ---- solute (dilute alloy, U field)
    p->D_l = c.num("composition/diffFluid", 0.0) * lu;
    p->D_s = c.num("composition/diffSolid", 0.0) * lu;
    p->k_part = c.num("composition/partition", 1.0);
    p->latent_C = c.num("solver/rejection", 0.5);
    // anti-trapping current of the dilute-alloy model (mkSolutalField::ATcurrent,
    // MaterKit/inc/mkSolutalField.h); Karma's coefficient 1/(2 sqrt 2) unless the case sets one
    p->at_coef = p->solute ? c.num("solver/antiTrapping", 0.5 / std::sqrt(2.0)) : 0.0;
    p->ref_c = p->solute ? c.num("composition/iniValue", 0.0) : 0.0;
    auto cf = c.nums("solver/faceSolute");
    if (cf.size() == 6)
        for (int k = 0; k < 6; k++) p->C_face[k] = cf[k];
    auto uin = c.nums("solver/inflow");
    if (uin.size() == 3)
        for (int k = 0; k < 3; k++) p->u_face[0][k] = uin[k];
    // rotating walls (bdyRotation / bdySpiral): angular velocity and the point on the axis
    auto wr = c.nums("solver/wallRotation");
    if (wr.size() == 3)
        for (int k = 0; k < 3; k++) p->wall_omega[k] = wr[k];
    auto wc = c.nums("solver/wallCenter");
    if (wc.size() == 3) {
        for (int k = 0; k < 3; k++) p->wall_center[k] = wc[k];
    } else {
        p->wall_center[0] = 0.5 * (p->nx - 1);
        p->wall_center[1] = 0.5 * (p->ny - 1);
        p->wall_center[2] = 0.5 * (p->nz - 1);
    }

    // ---- seeds / grains
    int nseed = 0;
    if (c.has("solver/seedsRandom")) {
        auto r = c.nums("solver/seedsRandom");
        nseed = r.size() > 0 ? (int)r[0] : 0;
    } else if (c.has("phasechange/seed/number")) {
        nseed = (int)c.num("phasechange/seed/number", 0);
    }
    if (nseed > MKG_MAX_GRAINS) { g_case_err = "too many seeds"; return -1; }
    p->ngrains = c.flag("solver/rotateGrains", false) ? nseed : 0;

    if (run_info) {
        run_info[0] = (int32_t)c.num("output/numLoops", 0);
        run_info[1] = (int32_t)c.num("output/numFrames", 0);
        run_info[2] = nseed;
        run_info[3] = 0;
    }
    return 0;
}

// Seeds of a case: explicit <phasechange><seed> entries or <solver><seedsRandom> count
// radius rngseed [margin].  Writes up to max seeds (centres 3n, radii n, rotations 9n row-major) and
// returns the count, or -1.
int mkg_seeds_from_xml(const char* path, int max, double* centers, double* radii, double* rot) {
    Case c;
    if (!load_case(path, c)) return -1;
    auto mesh = c.nums("geometry/MeshXYZ");
    if (mesh.size() != 3) { g_case_err = "geometry/MeshXYZ needs 3 values"; return -1; }
    if (c.has("solver/seedsRandom")) {
        auto r = c.nums("solver/seedsRandom");
        if (r.size() < 3) { g_case_err = "seedsRandom needs count radius rngseed"; return -1; }
        int n = (int)r[0];
        if (n > max) { g_case_err = "too many seeds"; return -1; }
        // optional fourth value: distance of the centres from every face (default: radius)
        const double margin = r.size() > 3 ? r[3] : r[1];
        std::vector<double> q(4 * (size_t)n);
        mkg_random_seeds((uint64_t)r[2], n, margin, (int)mesh[0], (int)mesh[1], (int)mesh[2],
                         centers, q.data());
        for (int k = 0; k < n; k++) {
            radii[k] = r[1];
            if (rot) mkg_quat_to_rot(&q[4 * k], rot + 9 * k);
        }
        return n;
    }
    int n = (int)c.num("phasechange/seed/number", 0);
    if (n > max) { g_case_err = "too many seeds"; return -1; }
    auto ctr = c.nums("phasechange/seed/center");
    auto ang = c.nums("phasechange/seed/angle");
    auto rad = c.nums("phasechange/seed/radius");
    if ((int)ctr.size() < 3 * n) { g_case_err = "seed/center needs 3 values per seed"; return -1; }
    for (int k = 0; k < n; k++) {
        for (int a = 0; a < 3; a++) centers[3 * k + a] = ctr[3 * k + a];
        radii[k] = rad.empty() ? 1.0 : rad[(size_t)k < rad.size() ? k : 0];
        if (rot) {
            double az = 0, ay = 0, ax = 0;
            if ((int)ang.size() >= 3 * (k + 1)) { az = ang[3 * k]; ay = ang[3 * k + 1]; ax = ang[3 * k + 2]; }
            euler_zyx_to_rot(az, ay, ax, rot + 9 * k);
        }
    }
    return n;
}

// Two-phase flow parameters from the <multiphases> block of a case file (the block of
// run_flow_multiphase's case, MaterKit/BoltzTechs - multiPhaseFlow.xml): fluidDense /
// fluidRare density and kinematicViscosity, W0, Mobility, sigma, force; kinematic viscosity
// and mobility are scaled by mk_dt / mk_dx^2 into lattice units, densities by the rare
// fluid's.  flow/acceleration is the body acceleration; solver/bcFlow the faces (periodic or
// bounce); solver/tauPhase the Cahn-Hilliard relaxation time (default 1).  The contact angle
// is not read (zero-flux faces give 90 degrees).  run_info: numLoops, numFrames, seeds.
int mkg_mp_params_from_xml(const char* path, mkg_mp_params* m, int32_t* run_info) {
    if (!path || !m) { g_case_err = "null argument"; return -1; }
    Case c;
    if (!load_case(path, c)) return -1;
    std::memset(m, 0, sizeof(*m));
    auto mesh = c.nums("geometry/MeshXYZ");
    if (mesh.size() != 3) { g_case_err = "geometry/MeshXYZ needs 3 values"; return -1; }
    m->nx = (int32_t)mesh[0];
    m->ny = (int32_t)mesh[1];
    m->nz = (int32_t)mesh[2];
    if (!c.has("multiphases")) { g_case_err = "the case has no <multiphases> block"; return -1; }
    const double dx = c.num("numerics/mk_dx", 1.0), dt = c.num("numerics/mk_dt", 1.0);
    const double lu = dt / (dx * dx);
    const double rr = c.num("multiphases/fluidRare/density", 1.0);
    if (!(rr > 0)) { g_case_err = "multiphases/fluidRare/density must be > 0"; return -1; }
    m->rho_l = c.num("multiphases/fluidDense/density", rr) / rr;
    m->rho_g = 1.0;
    m->nu_l = c.num("multiphases/fluidDense/kinematicViscosity", 0.1) * lu;
    m->nu_g = c.num("multiphases/fluidRare/kinematicViscosity", 0.1) * lu;
    m->width = c.num("multiphases/W0", 4.0);
    m->mobility = c.num("multiphases/Mobility", 0.1) * lu;
    m->sigma = c.num("multiphases/sigma", 0.001);
    auto f = c.nums("multiphases/force");
    if (f.size() == 3)
        for (int k = 0; k < 3; k++) m->force[k] = f[k];
    auto acc = c.nums("flow/acceleration");
    if (acc.size() == 3)
        for (int k = 0; k < 3; k++) m->gravity[k] = acc[k];
    m->tau_h = c.num("solver/tauPhase", 1.0);
    m->precision = c.word("solver/precision", "fp64") == "fp32" ? MKG_PREC_FP32 : MKG_PREC_FP64;
    int32_t bc[6] = {0, 0, 0, 0, 0, 0};
    if (!read_bc(c, "solver/bcFlow", bc)) return -1;
    for (int k = 0; k < 6; k++) m->bc[k] = bc[k];
    // contactAngle (degrees) makes every wall face a wetting face; 90 leaves them neutral
    m->contact_angle = c.num("multiphases/contactAngle", 90.0);
    if (!(m->contact_angle > 0.0 && m->contact_angle < 180.0)) {
        g_case_err = "multiphases/contactAngle must lie in (0, 180) degrees";
        return -1;
    }
    if (m->contact_angle != 90.0)
        for (int k = 0; k < 6; k++)
            if (bc[k] == MKG_BC_BOUNCE) m->wet_faces |= 1 << k;
    const std::string vm = c.word("solver/viscosityModel", "nu-linear");
    if (vm == "nu-linear") m->visc_model = MKG_MP_VISC_NU_LINEAR;
    else if (vm == "mu-linear") m->visc_model = MKG_MP_VISC_MU_LINEAR;
    else if (vm == "mu-harmonic") m->visc_model = MKG_MP_VISC_MU_HARMONIC;
    else {
        g_case_err = "solver/viscosityModel is nu-linear, mu-linear or mu-harmonic";
        return -1;
    }
    if (run_info) {
        run_info[0] = (int32_t)c.num("output/numLoops", 0);
        run_info[1] = (int32_t)c.num("output/numFrames", 1);
        run_info[2] = (int32_t)c.num("phasechange/seed/number", 0);
        run_info[3] = 0;
    }
    const std::string init = c.word("solver/twoPhaseInit", "layered");
    if (init != "layered" && init != "drops") {
        g_case_err = "solver/twoPhaseInit is 'layered' or 'drops'";
        return -1;
    }
    if (run_info) run_info[3] = init == "drops" ? 1 : 0;
    return 0;
}

}  // extern "C"
