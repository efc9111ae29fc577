// Implementation of the reference-API host layer (see mkshim.h).
#include "mkshim.h"

#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <sstream>

namespace {

void throw_io(const std::string& what) { throw mkt::mkIOException(what + ": " + mkg_last_error()); }

uint64_t splitmix(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
// counter-based uniform of one cell: splitmix64 of seed ^ (cell * K)
double cell_uniform(uint64_t seed, uint64_t cell) {
    uint64_t s = seed ^ (cell * 0xD1B54A32D192ED03ULL);
    return (double)(splitmix(&s) >> 11) * 0x1.0p-53;
}

// <solver><perturbation> amplitude seed </perturbation> of a case file (0 when absent).
bool read_perturbation(const std::string& xml, double& amp, uint64_t& seed) {
    std::ifstream f(xml);
    std::stringstream ss;
    ss << f.rdbuf();
    std::string s = ss.str();
    size_t a = s.find("<perturbation>"), b = s.find("</perturbation>");
    if (a == std::string::npos || b == std::string::npos) return false;
    std::istringstream is(s.substr(a + 14, b - a - 14));
    double sd = 0;
    if (!(is >> amp >> sd)) return false;
    seed = (uint64_t)sd;
    return true;
}

}  // namespace

void mkSpace::reset(const std::string& xmlFileName) {
    mkg_params p;
    if (mkg_params_from_xml(xmlFileName.c_str(), &p, nullptr)) throw_io("mkSpace");
    mx_ = p.nx;
    my_ = p.ny;
    mz_ = p.nz;
    xml_ = xmlFileName;
}

void readParameters(int& maxLoops, int& maxFrame, const std::string& xmlFileName) {
    mkg_params p;
    int32_t info[4];
    if (mkg_params_from_xml(xmlFileName.c_str(), &p, info)) throw_io("readParameters");
    maxLoops = info[0];
    maxFrame = info[1];
}

mkFluidsLBM::~mkFluidsLBM() {
    for (mkg_fvm_handle f : fvm_)
        if (f) mkg_fvm_destroy(f);
    if (h_) mkg_destroy(h_);
}

void mkFluidsLBM::fvm_evolve(int kind) {
    if (!h_) throw mkt::mkLogicException("evolution before initialize");
    const bool thermal = kind == kFvmThermal;
    if (!p_.phase_q || !(thermal ? p_.thermal : p_.solute))
        throw mkt::mkLogicException(thermal ? "evolution_thermal needs the phase and thermal fields"
                                            : "the finite-volume updates need the phase and "
                                              "solute fields");
    const size_t N = (size_t)p_.nx * p_.ny * p_.nz;
    if (fvm_live_ != kind) {
        const int fid = thermal ? MKG_FVM_TEMP : MKG_FVM_CONC;
        std::vector<double> phi(N), v(N);
        if (fvm_live_ >= 0 && (fvm_live_ == kFvmThermal) == thermal) {
            // the other solute handle holds the newest phi and c
            check(mkg_fvm_get(fvm_[fvm_live_], MKG_FVM_PHI, phi.data(), N));
            check(mkg_fvm_get(fvm_[fvm_live_], fid, v.data(), N));
        } else {
            fvm_flush();  // a handle of the other potential hands its fields to the lattices
            check(mkg_get_field(h_, MKG_F_PHI, phi.data(), MKG_DT_F64, N));
            check(mkg_get_field(h_, thermal ? MKG_F_TEMP : MKG_F_CONC, v.data(), MKG_DT_F64, N));
        }
        if (!fvm_[kind]) {
            mkg_params q = p_;
            q.ngrains = 0;
            q.thermal = thermal ? 1 : 0;
            if (kind != kFvmSoluteAT) q.at_coef = 0.0;
            check(mkg_fvm_create(&q, &fvm_[kind]));
        }
        check(mkg_fvm_set(fvm_[kind], MKG_FVM_PHI, phi.data(), N));
        check(mkg_fvm_set(fvm_[kind], fid, v.data(), N));
        fvm_live_ = kind;
    }
    check(mkg_fvm_step(fvm_[kind], 1, fvm_phase_next_ ? 1 : 0));
    fvm_phase_next_ = false;
}

void mkFluidsLBM::fvm_flush() const {
    if (fvm_live_ < 0) return;
    const bool thermal = fvm_live_ == kFvmThermal;
    const size_t N = (size_t)p_.nx * p_.ny * p_.nz;
    std::vector<double> phi(N), rho, u, T, c;
    check(mkg_fvm_get(fvm_[fvm_live_], MKG_FVM_PHI, phi.data(), N));
    if (p_.flow_q) {
        rho.resize(N);
        u.resize(3 * N);
        check(mkg_get_field(h_, MKG_F_RHO, rho.data(), MKG_DT_F64, N));
        check(mkg_get_field(h_, MKG_F_VEL, u.data(), MKG_DT_F64, 3 * N));
    }
    if (p_.thermal) {
        T.resize(N);
        if (thermal) check(mkg_fvm_get(fvm_[fvm_live_], MKG_FVM_TEMP, T.data(), N));
        else check(mkg_get_field(h_, MKG_F_TEMP, T.data(), MKG_DT_F64, N));
    }
    if (p_.solute) {
        c.resize(N);
        if (!thermal) check(mkg_fvm_get(fvm_[fvm_live_], MKG_FVM_CONC, c.data(), N));
        else check(mkg_get_field(h_, MKG_F_CONC, c.data(), MKG_DT_F64, N));
    }
    check(mkg_init_equilibrium(h_, p_.flow_q ? rho.data() : nullptr, p_.flow_q ? u.data() : nullptr,
                               p_.thermal ? T.data() : nullptr, p_.solute ? c.data() : nullptr,
                               phi.data()));
    fvm_live_ = -1;
}

void mkFluidsLBM::check(int rc) const {
    if (rc) throw mkt::mkLogicException(mkg_last_error());
}

void mkFluidsLBM::initialize(const std::string& xmlfname) {
    xml_ = xmlfname;
    if (mkg_params_from_xml(xmlfname.c_str(), &p_, run_info_)) throw_io("initialize");
    // seeds carry their rotations into the parameter block before the handle is created
    if (p_.phase_q && run_info_[2] > 0) {
        std::vector<double> c(3 * MKG_MAX_GRAINS), r(MKG_MAX_GRAINS), R(9 * MKG_MAX_GRAINS);
        int n = mkg_seeds_from_xml(xmlfname.c_str(), MKG_MAX_GRAINS, c.data(), r.data(), R.data());
        if (n < 0) throw_io("seeds");
        if (p_.ngrains)
            for (int k = 0; k < n; k++)
                for (int i = 0; i < 9; i++) p_.rot[k][i] = R[9 * k + i];
    }
    if (h_) mkg_destroy(h_);
    h_ = nullptr;
    check(mkg_create(&p_, &h_));
    // temperature: the case's initial value, or the conductive profile + noise of a
    // Rayleigh-Benard case
    const size_t N = (size_t)p_.nx * p_.ny * p_.nz;
    std::vector<double> T;
    if (p_.thermal) {
        T.assign(N, p_.T_ref);
        double amp = 0;
        uint64_t seed = 0;
        if (read_perturbation(xmlfname, amp, seed)) {
            const double Th = p_.T_face[4], Tc = p_.T_face[5];
            for (size_t id = 0; id < N; id++) {
                const int z = (int)(id % p_.nz);
                T[id] = Th - (Th - Tc) * (z + 0.5) / p_.nz + amp * (2.0 * cell_uniform(seed, id) - 1.0);
            }
        }
    }
    check(mkg_init_equilibrium(h_, nullptr, nullptr, p_.thermal ? T.data() : nullptr, nullptr,
                               nullptr));
}

void mkFluidsLBM::initialize_geometry() {
    if (!h_) throw mkt::mkLogicException("initialize_geometry before initialize");
    if (!(p_.phase_q && run_info_[2] > 0)) return;
    std::vector<double> c(3 * MKG_MAX_GRAINS), r(MKG_MAX_GRAINS);
    int n = mkg_seeds_from_xml(xml_.c_str(), MKG_MAX_GRAINS, c.data(), r.data(), nullptr);
    if (n < 0) throw_io("seeds");
    check(mkg_init_seeds(h_, n, c.data(), r.data()));
    if (p_.solute) {
        // c = A(phi) U0 in every cell, U0 = the case's initial supersaturation variable
        const size_t N = (size_t)p_.nx * p_.ny * p_.nz;
        std::vector<double> phi(N), C(N);
        check(mkg_get_field(h_, MKG_F_PHI, phi.data(), MKG_DT_F64, N));
        for (size_t i = 0; i < N; i++)
            C[i] = 0.5 * ((1 + p_.k_part) - (1 - p_.k_part) * phi[i]) * p_.ref_c;
        std::vector<double> T;
        check(mkg_init_equilibrium(h_, nullptr, nullptr, nullptr, C.data(), nullptr));
    }
}

void mkFluidsLBM::advance(int mask) {
    if (!h_) throw mkt::mkLogicException("relaxation before initialize");
    fvm_flush();
    check(mkg_step(h_, 1, mask));
    pending_ &= ~mask;
}

void mkFluidsLBM::step(int nsteps) {
    if (!h_) throw mkt::mkLogicException("step before initialize");
    fvm_flush();
    check(mkg_step(h_, nsteps, MKG_FLOW | MKG_PHASE | MKG_THERMAL | MKG_SOLUTE));
    pending_ = 0;
}

std::vector<double> mkFluidsLBM::field(int id) const {
    const size_t N = (size_t)p_.nx * p_.ny * p_.nz;
    std::vector<double> v(id == MKG_F_VEL ? 3 * N : N);
    fvm_flush();
    check(mkg_get_field(h_, id, v.data(), MKG_DT_F64, v.size()));
    return v;
}

mkNode* mkFluidsLBM::fetch_nodes() {
    if (!h_) throw mkt::mkLogicException("fetch_nodes before initialize (or on a class without "
                                         "the lattice handle)");
    fvm_flush();
    const size_t N = (size_t)p_.nx * p_.ny * p_.nz;
    nodes_.resize(N);
    auto pops = [&](int id, int q, mkPDFs mkNode::*member) {
        if (!q) return;
        std::vector<double> v((size_t)q * N);
        check(mkg_get_field(h_, id, v.data(), MKG_DT_F64, v.size()));
        for (size_t k = 0; k < N; k++) {
            mkPDFs& d = nodes_[k].*member;
            d.resize(q);
            for (int i = 0; i < q; i++) d[i] = v[(size_t)i * N + k];
        }
    };
    pops(MKG_F_POP_FLOW, p_.flow_q, &mkNode::f_hlp);
    pops(MKG_F_POP_PHASE, p_.phase_q, &mkNode::g_hlp);
    pops(MKG_F_POP_THERMAL, p_.thermal ? 7 : 0, &mkNode::h_hlp);
    pops(MKG_F_POP_SOLUTE, p_.solute ? 7 : 0, &mkNode::x_hlp);
    auto scalar = [&](int id, bool present, auto set) {
        if (!present) return;
        std::vector<double> v(N);
        check(mkg_get_field(h_, id, v.data(), MKG_DT_F64, N));
        for (size_t k = 0; k < N; k++) set(nodes_[k], v[k]);
    };
    scalar(MKG_F_RHO, p_.flow_q != 0, [](mkNode& n, double v) { n.density = v; });
    scalar(MKG_F_TEMP, p_.thermal != 0, [](mkNode& n, double v) { n.temperature = v; });
    scalar(MKG_F_CONC, p_.solute != 0, [](mkNode& n, double v) { n.X_solutes.assign(1, v); });
    if (p_.phase_q) {
        std::vector<double> now(N), prev(N);
        check(mkg_get_field(h_, MKG_F_PHI, now.data(), MKG_DT_F64, N));
        check(mkg_get_field(h_, MKG_F_PHI_PREV, prev.data(), MKG_DT_F64, N));
        for (size_t k = 0; k < N; k++) {
            nodes_[k].phi = now[k];
            nodes_[k].d_phi = now[k] - prev[k];
        }
    }
    if (p_.flow_q) {
        std::vector<double> u(3 * N);
        check(mkg_get_field(h_, MKG_F_VEL, u.data(), MKG_DT_F64, u.size()));
        for (size_t k = 0; k < N; k++) nodes_[k].vel.set(u[3 * k], u[3 * k + 1], u[3 * k + 2]);
    }
    std::vector<int8_t> fl(N);
    check(mkg_get_field(h_, MKG_F_FLAG, fl.data(), MKG_DT_I8, N));
    std::vector<int16_t> gr(N);
    check(mkg_get_field(h_, MKG_F_GRAIN, gr.data(), MKG_DT_I16, N));
    for (size_t k = 0; k < N; k++) {
        nodes_[k].flag = fl[k];
        nodes_[k].grain_id = gr[k];
    }
    node = nodes_.data();
    return node;
}

void mkFluidsLBM::store_nodes() {
    if (!h_ || nodes_.empty()) throw mkt::mkLogicException("store_nodes before fetch_nodes");
    fvm_flush();
    const size_t N = nodes_.size();
    auto pops = [&](int id, int q, mkPDFs mkNode::*member) {
        if (!q) return;
        std::vector<double> v((size_t)q * N);
        for (size_t k = 0; k < N; k++)
            for (int i = 0; i < q; i++) v[(size_t)i * N + k] = (nodes_[k].*member)[i];
        check(mkg_set_field(h_, id, v.data(), MKG_DT_F64, v.size()));
    };
    pops(MKG_F_POP_FLOW, p_.flow_q, &mkNode::f_hlp);
    pops(MKG_F_POP_PHASE, p_.phase_q, &mkNode::g_hlp);
    pops(MKG_F_POP_THERMAL, p_.thermal ? 7 : 0, &mkNode::h_hlp);
    pops(MKG_F_POP_SOLUTE, p_.solute ? 7 : 0, &mkNode::x_hlp);
    std::vector<int8_t> fl(N);
    for (size_t k = 0; k < N; k++) fl[k] = (int8_t)nodes_[k].flag;
    check(mkg_set_field(h_, MKG_F_FLAG, fl.data(), MKG_DT_I8, N));
}

double mkFluidsLBM::Thermal_NusseltRBC() {
    fvm_flush();
    double v[4];
    check(mkg_reduce(h_, MKG_R_NUSSELT, v));
    return v[0];
}

double mkFluidsLBM::Solutal_NusseltRBC() {
    // solutal analogue on the same definition: 1 + <u_z c> H / (D dC)
    const size_t N = (size_t)p_.nx * p_.ny * p_.nz;
    std::vector<double> u = field(MKG_F_VEL), c = field(MKG_F_CONC);
    double s = 0;
    for (size_t i = 0; i < N; i++) s += u[3 * i + 2] * c[i];
    const double dC = p_.C_face[4] - p_.C_face[5];
    if (dC == 0 || p_.D_l == 0) return 1.0;
    return 1.0 + (s / N) * p_.nz / (p_.D_l * dC);
}

// Mean velocity in the mid cross-section normal to nPlane (MaterKit/inc/mkFluidsLBM.h:54-55).
void mkFluidsLBM::check_variables(int loops, const char nPlane) {
    if (!p_.flow_q) return;
    std::vector<double> u = field(MKG_F_VEL);
    double s[3] = {0, 0, 0};
    long cnt = 0;
    for (int x = 0; x < p_.nx; x++)
        for (int y = 0; y < p_.ny; y++)
            for (int z = 0; z < p_.nz; z++) {
                bool in = (nPlane == 'x' && x == p_.nx / 2) || (nPlane == 'y' && y == p_.ny / 2) ||
                          (nPlane == 'z' && z == p_.nz / 2);
                if (!in) continue;
                size_t id = ((size_t)x * p_.ny + y) * p_.nz + z;
                for (int k = 0; k < 3; k++) s[k] += u[3 * id + k];
                cnt++;
            }
    if (cnt)
        std::printf("loops %d  mean velocity in the %c mid-plane: %.10e %.10e %.10e\n", loops, nPlane,
                    s[0] / cnt, s[1] / cnt, s[2] / cnt);
}

void mkFluidsLBM::read_parameters(const std::string& xmlfname) {
    xml_ = xmlfname;
    if (mkg_params_from_xml(xmlfname.c_str(), &p_, run_info_)) throw_io("read_parameters");
}

// Fields of the frame writers: name and values (N each; velocity as three components).
static std::vector<std::pair<std::string, std::vector<double>>> frame_fields(
    const mkFluidsLBM& f, const mkg_params& p) {
    std::vector<std::pair<std::string, std::vector<double>>> out;
    const size_t N = (size_t)p.nx * p.ny * p.nz;
    if (p.flow_q) {
        std::vector<double> r = f.field(MKG_F_RHO);
        for (double& v : r) v -= 1.0;
        out.emplace_back("density_difference", std::move(r));
    }
    if (p.phase_q) out.emplace_back("order_parameter", f.field(MKG_F_PHI));
    if (p.solute) out.emplace_back("X0_solute", f.field(MKG_F_CONC));
    if (p.thermal) out.emplace_back("temperature", f.field(MKG_F_TEMP));
    if (p.flow_q) {
        std::vector<double> u = f.field(MKG_F_VEL), c[3];
        for (int k = 0; k < 3; k++) {
            c[k].resize(N);
            for (size_t i = 0; i < N; i++) c[k][i] = u[3 * i + k];
        }
        out.emplace_back("u", std::move(c[0]));
        out.emplace_back("v", std::move(c[1]));
        out.emplace_back("w", std::move(c[2]));
    }
    return out;
}

// Tecplot ASCII zone over the box [lo, hi) of cells, x fastest.
static void write_plt(const std::string& path, const mkg_params& p,
                      const std::vector<std::pair<std::string, std::vector<double>>>& fields,
                      const int (&lo)[3], const int (&hi)[3]) {
    std::ofstream o(path);
    if (!o) throw mkt::mkIOException("cannot write " + path);
    o << "TITLE = \"fluid_state\"\nVARIABLES = \"x\", \"y\", \"z\"";
    for (const auto& f : fields) o << ", \"" << f.first << "\"";
    o << "\nZONE I=" << hi[0] - lo[0] << ", J=" << hi[1] - lo[1] << ", K=" << hi[2] - lo[2]
      << ", F=POINT\n" << std::setprecision(9);
    for (int z = lo[2]; z < hi[2]; z++)
        for (int y = lo[1]; y < hi[1]; y++)
            for (int x = lo[0]; x < hi[0]; x++) {
                const size_t id = ((size_t)x * p.ny + y) * p.nz + z;
                o << x + 0.5 << " " << y + 0.5 << " " << z + 0.5;
                for (const auto& f : fields) o << " " << f.second[id];
                o << "\n";
            }
}

void mkFluidsLBM::write_pltfile(int frame) {
    std::error_code ec;
    std::filesystem::create_directories(workspace_ + "plt", ec);
    const int lo[3] = {0, 0, 0}, hi[3] = {p_.nx, p_.ny, p_.nz};
    write_plt(workspace_ + "plt/fluid_" + std::to_string(frame) + ".plt", p_, frame_fields(*this, p_),
              lo, hi);
}

void mkFluidsLBM::write_pltfile_2D(int loops, char nPlane) {
    const int ax = nPlane == 'x' ? 0 : (nPlane == 'y' ? 1 : (nPlane == 'z' ? 2 : -1));
    if (ax < 0) throw mkt::mkLogicException("write_pltfile_2D: nPlane is 'x', 'y' or 'z'");
    std::error_code ec;
    std::filesystem::create_directories(workspace_ + "plt", ec);
    int lo[3] = {0, 0, 0}, hi[3] = {p_.nx, p_.ny, p_.nz};
    lo[ax] = hi[ax] / 2;
    hi[ax] = lo[ax] + 1;
    write_plt(workspace_ + "plt/fluid2D_" + std::string(1, nPlane) + "_t" + std::to_string(loops) +
                  ".plt",
              p_, frame_fields(*this, p_), lo, hi);
}

void mkFluidsLBM::write_vtkfile_2D(int loops) {
    const int Mx = p_.nx, My = p_.ny, z0 = p_.nz / 2;
    std::error_code ec;
    std::filesystem::create_directories(workspace_ + "vtk", ec);
    const std::string fn = workspace_ + "vtk/fluid2D_t" + std::to_string(loops) + ".vtk";
    std::ofstream o(fn);
    if (!o) throw mkt::mkIOException("cannot write " + fn);
    o << "# vtk DataFile Version 3.0\nfluid_state\nASCII\nDATASET RECTILINEAR_GRID\n";
    o << "DIMENSIONS " << Mx << " " << My << " 1\n";
    o << "X_COORDINATES " << Mx << " float\n";
    for (int i = 0; i < Mx; i++) o << i + 0.5 << " ";
    o << "\nY_COORDINATES " << My << " float\n";
    for (int i = 0; i < My; i++) o << i + 0.5 << " ";
    o << "\nZ_COORDINATES 1 float\n" << z0 + 0.5 << "\n";
    o << "POINT_DATA " << (size_t)Mx * My << "\n" << std::setprecision(9);
    for (const auto& f : frame_fields(*this, p_)) {
        o << "SCALARS " << f.first << " float 1\nLOOKUP_TABLE default\n";
        for (int y = 0; y < My; y++)
            for (int x = 0; x < Mx; x++) o << f.second[((size_t)x * My + y) * p_.nz + z0] << "\n";
    }
}

void mkFluidsLBM::write_vtkfile(int loops) {
    const int Mx = p_.nx, My = p_.ny, Mz = p_.nz;
    const size_t N = (size_t)Mx * My * Mz;
    std::ostringstream fn;
    std::error_code ec;
    std::filesystem::create_directories(workspace_ + "vtk", ec);
    fn << workspace_ << "vtk/fluid_t" << loops << ".vtk";
    std::ofstream o(fn.str());
    if (!o) throw mkt::mkIOException("cannot write " + fn.str());
    o << "# vtk DataFile Version 3.0\nfluid_state\nASCII\nDATASET RECTILINEAR_GRID\n";
    o << "DIMENSIONS " << Mx << " " << My << " " << Mz << "\n";
    auto coords = [&](const char* n, int m) {
        o << n << " " << m << " float\n";
        for (int i = 0; i < m; i++) o << i + 0.5 << " ";
        o << "\n";
    };
    coords("X_COORDINATES", Mx);
    coords("Y_COORDINATES", My);
    coords("Z_COORDINATES", Mz);
    o << "POINT_DATA " << N << "\n";
    o << std::setprecision(9);
    auto scalar = [&](const char* name, const std::vector<double>& v, double off) {
        o << "SCALARS " << name << " float 1\nLOOKUP_TABLE default\n";
        for (int z = 0; z < Mz; z++)
            for (int y = 0; y < My; y++)
                for (int x = 0; x < Mx; x++) o << v[((size_t)x * My + y) * Mz + z] - off << "\n";
    };
    if (p_.flow_q) scalar("density_difference", field(MKG_F_RHO), 1.0);
    if (p_.phase_q) {
        std::vector<int16_t> g(N);
        fvm_flush();
        check(mkg_get_field(h_, MKG_F_GRAIN, g.data(), MKG_DT_I16, N));
        std::vector<double> gd(g.begin(), g.end());
        scalar("grain_id", gd, 0.0);
        std::vector<double> phi = field(MKG_F_PHI), prev = field(MKG_F_PHI_PREV);
        scalar("order_parameter", phi, 0.0);
        for (size_t i = 0; i < N; i++) prev[i] = phi[i] - prev[i];
        scalar("delta_order_parameter", prev, 0.0);
    }
    if (p_.solute) scalar("X0_solute", field(MKG_F_CONC), 0.0);
    if (p_.thermal) scalar("temperature", field(MKG_F_TEMP), 0.0);
    if (p_.flow_q) {
        std::vector<double> u = field(MKG_F_VEL);
        o << "VECTORS velocity_vector float\n";
        for (int z = 0; z < Mz; z++)
            for (int y = 0; y < My; y++)
                for (int x = 0; x < Mx; x++) {
                    size_t id = ((size_t)x * My + y) * Mz + z;
                    o << u[3 * id] << " " << u[3 * id + 1] << " " << u[3 * id + 2] << "\n";
                }
    }
}

// ------------------------------------------------------------------ driver utilities
mkCPUTimer* timer = nullptr;
namespace mkt {
std::string paraXmlFileName;
}
const std::string workspace = "./";

mkCPUTimer::mkCPUTimer(const std::string label) : label_(label), t0_(std::chrono::steady_clock::now()) {}

mkCPUTimer::~mkCPUTimer() { std::printf("%s: %.3f s\n", label_.c_str(), GetTime(0)); }

double mkCPUTimer::GetTime(int /*flag*/) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

double mkCPUTimer::GetSystemTime() {
    return std::chrono::duration<double>(std::chrono::system_clock::now().time_since_epoch())
        .count();
}

extern "C" int make_datafolder() {
    std::error_code ec;
    std::filesystem::create_directories(workspace + "vtk", ec);
    return ec ? -1 : 0;
}

extern "C" int make_numthreads() { return 1; }

void mkFluidsMultiphase::update_geometry(mkFluidsMRT& /*dendrite*/) {
    throw mkt::mkLogicException(
        "mkFluidsMultiphase::update_geometry (two-phase flow around a growing dendrite) is not "
        "provided");
}

// ------------------------------------------------------------------ anisotropic multi-grain field
mkMultiphaseField::~mkMultiphaseField() {
    if (m_) mkg_mpf_destroy(m_);
}

void mkMultiphaseField::initialize(const std::string& xmlfname) {
    xml_ = xmlfname;
    if (mkg_params_from_xml(xmlfname.c_str(), &p_, run_info_)) throw_io("initialize");
    seed_number = run_info_[2];
    phase_number = seed_number;
    if (phase_number < 1 || phase_number > MKG_MPF_MAX)
        throw mkt::mkLogicException("the multi-grain phase field needs 1.." +
                                    std::to_string(MKG_MPF_MAX) + " seeds (one grain each)");
    std::vector<double> c(3 * MKG_MAX_GRAINS), r(MKG_MAX_GRAINS), R(9 * MKG_MAX_GRAINS);
    if (mkg_seeds_from_xml(xmlfname.c_str(), MKG_MAX_GRAINS, c.data(), r.data(), R.data()) < 0)
        throw_io("seeds");
    vec_center.assign(c.begin(), c.begin() + 3 * phase_number);
    vec_angle.assign(R.begin(), R.begin() + 9 * phase_number);  // as rotation matrices
    check(mkg_mpf_params_default(&mp_));
    mp_.nx = p_.nx;
    mp_.ny = p_.ny;
    mp_.nz = p_.nz;
    mp_.n = phase_number;
    mp_.precision = p_.precision;
    mp_.model = MKG_MPF_KARMA_RAPPEL;
    mp_.thermal = p_.thermal;
    mp_.solute = p_.solute;
    mp_.coupling = p_.pf_coupling;
    mp_.dt = p_.pf_dt;
    mp_.wdx = std::sqrt(p_.pf_wdx2);
    mp_.lambda = p_.lambda;
    mp_.alpha = p_.alpha_l;
    mp_.latent_T = p_.latent_T;
    mp_.T_ref = p_.T_ref;
    mp_.D_l = p_.D_l;
    mp_.D_s = p_.D_s;
    mp_.k_part = p_.k_part;
    mp_.q_min = p_.q_min;
    mp_.latent_C = p_.latent_C;
    mp_.at_coef = p_.at_coef;
    for (int f = 0; f < 6; f++) {
        mp_.bc[f] = p_.bc_g[f];
        mp_.bc_h[f] = p_.bc_h[f];
        mp_.bc_c[f] = p_.bc_c[f];
        mp_.T_face[f] = p_.T_face[f];
        mp_.C_face[f] = p_.C_face[f];
    }
    for (int j = 0; j < phase_number; j++) {
        mp_.u[j][0] = 0.0;  // grains do not move
        mp_.kappa[j] = 1.0;
        mp_.aniso_kind[j] = p_.aniso_kind;
        mp_.eps4[j] = p_.eps4;
        mp_.aniso[j][0] = p_.aniso[0];
        mp_.aniso[j][1] = p_.aniso[1];
        for (int k = 0; k < 9; k++) mp_.rot[j][k] = R[9 * j + k];
    }
    if (m_) mkg_mpf_destroy(m_);
    m_ = nullptr;
    check(mkg_mpf_create(&mp_, &m_));
}

void mkMultiphaseField::initialize_geometry() {
    if (!m_) throw mkt::mkLogicException("initialize_geometry before initialize");
    // phi_j = -tanh((|x - c_j| - r_j) / (sqrt 2 W0/dx)); T = the case's initial value;
    // c = A(phi) U0 of the combined phi, U0 = the case's initial supersaturation variable
    std::vector<double> c(3 * MKG_MAX_GRAINS), r(MKG_MAX_GRAINS);
    if (mkg_seeds_from_xml(xml_.c_str(), MKG_MAX_GRAINS, c.data(), r.data(), nullptr) < 0)
        throw_io("seeds");
    const size_t N = (size_t)mp_.nx * mp_.ny * mp_.nz;
    const int n = mp_.n;
    std::vector<double> phi((size_t)n * N), C(N);
    const double width = std::sqrt(2.0) * mp_.wdx;
    for (int j = 0; j < n; j++)
        for (int x = 0; x < mp_.nx; x++)
            for (int y = 0; y < mp_.ny; y++)
                for (int z = 0; z < mp_.nz; z++) {
                    const double dx = x - c[3 * j], dy = y - c[3 * j + 1], dz = z - c[3 * j + 2];
                    const double d = std::sqrt(dx * dx + dy * dy + dz * dz) - r[j];
                    phi[(size_t)j * N + ((size_t)x * mp_.ny + y) * mp_.nz + z] = -std::tanh(d / width);
                }
    for (size_t i = 0; i < N; i++) {
        double comb = n - 1.0;
        for (int j = 0; j < n; j++) comb += phi[(size_t)j * N + i];
        comb = std::fmin(1.0, std::fmax(-1.0, comb));
        C[i] = 0.5 * ((1 + mp_.k_part) - (1 - mp_.k_part) * comb) * p_.ref_c;
    }
    check(mkg_mpf_init_fields(m_, phi.data(), nullptr, mp_.solute ? C.data() : nullptr));
}

void mkMultiphaseField::advance(int mask) {
    if (!m_) throw mkt::mkLogicException("relaxation before initialize");
    if (mask & (MKG_PHASE | MKG_FLOW)) check(mkg_mpf_step(m_, 1));
    pending_ &= ~mask;
}

void mkMultiphaseField::driven(int coupling, const char* name) {
    if (!m_) throw mkt::mkLogicException(std::string(name) + " before initialize");
    if (mp_.coupling != coupling)
        throw mkt::mkLogicException(std::string(name) + ": the case drives the grains by " +
                                    (mp_.coupling == 1 ? "T" : (mp_.coupling == 2 ? "U" : "nothing")));
    advance(MKG_PHASE);
}

void mkMultiphaseField::relaxation_AllenCahn() {
    throw mkt::mkLogicException(
        "mkMultiphaseField::relaxation_AllenCahn is not provided: the Allen-Cahn order "
        "parameters are mkFluidsUser's (mkg_mpf model MKG_MPF_TEMPLATE)");
}

void mkMultiphaseField::evolution_anisotropy() {
    throw mkt::mkLogicException(
        "mkMultiphaseField::evolution_anisotropy (finite volumes) is not provided: the grains "
        "evolve on the lattice (relaxation_anisotropy_thermal / relaxation_solute)");
}

bool mkMultiphaseField::grain_query(int kind, int id, mkVector3d& N_v, mkReal& A_s,
                                    const mkVector3d& normal) {
    if (id < 1 || id > mp_.n)
        throw mkt::mkLogicException("anisotropy: grain id must be 1.." + std::to_string(mp_.n));
    const int j = id - 1;
    const int rc = mkg_aniso_eval(kind, mp_.eps4[j], mp_.aniso[j], mp_.rot[j], normal.ptr(),
                                  &A_s, N_v.ptr());
    if (rc < 0) throw mkt::mkLogicException(mkg_last_error());
    return rc == 0;
}

std::vector<double> mkMultiphaseField::order_parameters() const {
    const size_t N = (size_t)mp_.nx * mp_.ny * mp_.nz;
    std::vector<double> v((size_t)(mp_.n + 1) * N);
    check(mkg_mpf_get(m_, MKG_MPF_PHI, v.data(), v.size()));
    return v;
}

void mkMultiphaseField::write_vtkfile(int loops) {
    const int Mx = mp_.nx, My = mp_.ny, Mz = mp_.nz;
    const size_t N = (size_t)Mx * My * Mz;
    std::ostringstream fn;
    std::error_code ec;
    std::filesystem::create_directories(workspace_ + "vtk", ec);
    fn << workspace_ << "vtk/grains_t" << loops << ".vtk";
    std::ofstream o(fn.str());
    if (!o) throw mkt::mkIOException("cannot write " + fn.str());
    o << "# vtk DataFile Version 3.0\ngrains\nASCII\nDATASET RECTILINEAR_GRID\n";
    o << "DIMENSIONS " << Mx << " " << My << " " << Mz << "\n";
    auto coords = [&](const char* nm, int m) {
        o << nm << " " << m << " float\n";
        for (int i = 0; i < m; i++) o << i + 0.5 << " ";
        o << "\n";
    };
    coords("X_COORDINATES", Mx);
    coords("Y_COORDINATES", My);
    coords("Z_COORDINATES", Mz);
    o << "POINT_DATA " << N << "\n" << std::setprecision(9);
    const std::vector<double> ph = order_parameters();
    auto scalar = [&](const std::string& nm, const double* v) {
        o << "SCALARS " << nm << " float 1\nLOOKUP_TABLE default\n";
        for (int z = 0; z < Mz; z++)
            for (int y = 0; y < My; y++)
                for (int x = 0; x < Mx; x++) o << v[((size_t)x * My + y) * Mz + z] << "\n";
    };
    for (int j = 0; j < mp_.n; j++) scalar("phi_" + std::to_string(j), ph.data() + (size_t)j * N);
    scalar("phi", ph.data() + (size_t)mp_.n * N);
    // grain id: the order parameter with the largest phi where any exceeds 0, else 0
    std::vector<double> gid(N, 0.0);
    for (size_t i = 0; i < N; i++) {
        double best = 0.0;
        for (int j = 0; j < mp_.n; j++)
            if (ph[(size_t)j * N + i] > best) {
                best = ph[(size_t)j * N + i];
                gid[i] = j + 1;
            }
    }
    scalar("grain", gid.data());
}

// ------------------------------------------------------------------ multi-phase field
mkFluidsUser::~mkFluidsUser() {
    if (m_) mkg_mpf_destroy(m_);
}

void mkFluidsUser::initialize(const std::string& xmlfname) {
    xml_ = xmlfname;
    if (mkg_params_from_xml(xmlfname.c_str(), &p_, run_info_)) throw_io("initialize");
    const int n = run_info_[2];
    if (n < 1 || n > MKG_MPF_MAX)
        throw mkt::mkLogicException("the multi-phase field needs 1.." + std::to_string(MKG_MPF_MAX) +
                                    " seeds (one order parameter each)");
    check(mkg_mpf_params_default(&mp_));
    mp_.nx = p_.nx;
    mp_.ny = p_.ny;
    mp_.nz = p_.nz;
    mp_.n = n;
    mp_.precision = p_.precision;
    mp_.thermal = p_.thermal;
    mp_.dt = p_.pf_dt;
    mp_.wdx = std::sqrt(p_.pf_wdx2);
    mp_.alpha = p_.alpha_l;
    mp_.latent_T = p_.latent_T;
    mp_.T_ref = p_.T_ref;
    for (int f = 0; f < 6; f++) {
        mp_.bc[f] = p_.bc_g[f];
        mp_.bc_h[f] = p_.bc_h[f];
        mp_.T_face[f] = p_.T_face[f];
    }
    if (m_) mkg_mpf_destroy(m_);
    m_ = nullptr;
    check(mkg_mpf_create(&mp_, &m_));
}

void mkFluidsUser::initialize_geometry() {
    if (!m_) throw mkt::mkLogicException("initialize_geometry before initialize");
    std::vector<double> c(3 * MKG_MAX_GRAINS), r(MKG_MAX_GRAINS);
    const int n = mkg_seeds_from_xml(xml_.c_str(), MKG_MAX_GRAINS, c.data(), r.data(), nullptr);
    if (n < 0) throw_io("seeds");
    // phi_j = -tanh((|x - c_j| - r_j) / (sqrt 2 W0/dx)): one order parameter per seed
    const size_t N = (size_t)mp_.nx * mp_.ny * mp_.nz;
    std::vector<double> phi((size_t)mp_.n * N);
    const double width = std::sqrt(2.0) * mp_.wdx;
    for (int j = 0; j < mp_.n; j++)
        for (int x = 0; x < mp_.nx; x++)
            for (int y = 0; y < mp_.ny; y++)
                for (int z = 0; z < mp_.nz; z++) {
                    const double dx = x - c[3 * j], dy = y - c[3 * j + 1], dz = z - c[3 * j + 2];
                    const double d = std::sqrt(dx * dx + dy * dy + dz * dz) - r[j];
                    phi[(size_t)j * N + ((size_t)x * mp_.ny + y) * mp_.nz + z] = -std::tanh(d / width);
                }
    check(mkg_mpf_init(m_, phi.data(), nullptr));
}

void mkFluidsUser::advance(int mask) {
    if (!m_) throw mkt::mkLogicException("relaxation before initialize");
    // the thermal lattice steps with the order parameters
    if (mask & (MKG_PHASE | MKG_FLOW)) check(mkg_mpf_step(m_, 1));
    pending_ &= ~mask;
}

std::vector<double> mkFluidsUser::order_parameters() const {
    const size_t N = (size_t)mp_.nx * mp_.ny * mp_.nz;
    std::vector<double> v((size_t)(mp_.n + 1) * N);
    check(mkg_mpf_get(m_, MKG_MPF_PHI, v.data(), v.size()));
    return v;
}

void mkFluidsUser::write_vtkfile(int loops) {
    const int Mx = mp_.nx, My = mp_.ny, Mz = mp_.nz;
    const size_t N = (size_t)Mx * My * Mz;
    std::ostringstream fn;
    std::error_code ec;
    std::filesystem::create_directories(workspace_ + "vtk", ec);
    fn << workspace_ << "vtk/fluid_t" << loops << ".vtk";
    std::ofstream o(fn.str());
    if (!o) throw mkt::mkIOException("cannot write " + fn.str());
    o << "# vtk DataFile Version 3.0\nfluid_state\nASCII\nDATASET RECTILINEAR_GRID\n";
    o << "DIMENSIONS " << Mx << " " << My << " " << Mz << "\n";
    auto coords = [&](const char* nm, int m) {
        o << nm << " " << m << " float\n";
        for (int i = 0; i < m; i++) o << i + 0.5 << " ";
        o << "\n";
    };
    coords("X_COORDINATES", Mx);
    coords("Y_COORDINATES", My);
    coords("Z_COORDINATES", Mz);
    o << "POINT_DATA " << N << "\n" << std::setprecision(9);
    auto scalar = [&](const std::string& name, const double* v) {
        o << "SCALARS " << name << " float 1\nLOOKUP_TABLE default\n";
        for (int z = 0; z < Mz; z++)
            for (int y = 0; y < My; y++)
                for (int x = 0; x < Mx; x++) o << v[((size_t)x * My + y) * Mz + z] << "\n";
    };
    const std::vector<double> ph = order_parameters();
    scalar("order_parameter", ph.data() + (size_t)mp_.n * N);
    for (int j = 0; j < mp_.n; j++)
        scalar("order_parameter_" + std::to_string(j), ph.data() + (size_t)j * N);
    if (mp_.thermal) {
        std::vector<double> h(7 * N), T(N, 0.0);
        check(mkg_mpf_get(m_, MKG_MPF_THERMAL, h.data(), h.size()));
        for (int i = 0; i < 7; i++)
            for (size_t k = 0; k < N; k++) T[k] += h[(size_t)i * N + k];
        scalar("temperature", T.data());
    }
}

// ------------------------------------------------------------------ two-phase flow
mkFluidsMultiphase::~mkFluidsMultiphase() {
    if (m_) mkg_mp_destroy(m_);
}

void mkFluidsMultiphase::initialize(const std::string& xmlfname) {
    xml_ = xmlfname;
    if (mkg_mp_params_from_xml(xmlfname.c_str(), &mp_, run_info_)) throw_io("initialize");
    if (m_) mkg_mp_destroy(m_);
    m_ = nullptr;
    check(mkg_mp_create(&mp_, &m_));
}

void mkFluidsMultiphase::initialize_geometry() {
    if (!m_) throw mkt::mkLogicException("initialize_geometry before initialize");
    const size_t N = (size_t)mp_.nx * mp_.ny * mp_.nz;
    std::vector<double> phi(N);
    const bool drops = run_info_[3] == 1;
    const double W = mp_.width;
    if (drops) {
        std::vector<double> c(3 * MKG_MAX_GRAINS), r(MKG_MAX_GRAINS);
        const int n = mkg_seeds_from_xml(xml_.c_str(), MKG_MAX_GRAINS, c.data(), r.data(), nullptr);
        if (n < 0) throw_io("seeds");
        for (int x = 0; x < mp_.nx; x++)
            for (int y = 0; y < mp_.ny; y++)
                for (int z = 0; z < mp_.nz; z++) {
                    double best = 1e300;
                    for (int k = 0; k < n; k++) {
                        const double dx = x - c[3 * k], dy = y - c[3 * k + 1], dz = z - c[3 * k + 2];
                        best = std::min(best, std::sqrt(dx * dx + dy * dy + dz * dz) - r[k]);
                    }
                    phi[((size_t)x * mp_.ny + y) * mp_.nz + z] = 0.5 - 0.5 * std::tanh(2.0 * best / W);
                }
    } else {
        for (int x = 0; x < mp_.nx; x++)
            for (int y = 0; y < mp_.ny; y++)
                for (int z = 0; z < mp_.nz; z++)
                    phi[((size_t)x * mp_.ny + y) * mp_.nz + z] =
                        0.5 - 0.5 * std::tanh(2.0 * (y + 0.5 - 0.5 * mp_.ny) / W);
    }
    check(mkg_mp_init(m_, phi.data(), nullptr, nullptr));
}

void mkFluidsMultiphase::initialize_wetting(const std::string& xmlfname) {
    xml_ = xmlfname;
    if (mkg_mp_params_from_xml(xmlfname.c_str(), &mp_, run_info_)) throw_io("initialize_wetting");
    bool walls = false;
    for (int f = 0; f < 6; f++) walls |= mp_.bc[f] == MKG_BC_BOUNCE;
    if (!walls) mp_.bc[2] = mp_.bc[3] = MKG_BC_BOUNCE;
    mp_.wet_faces = 0;
    if (mp_.contact_angle != 90.0)
        for (int f = 0; f < 6; f++)
            if (mp_.bc[f] == MKG_BC_BOUNCE) mp_.wet_faces |= 1 << f;
    if (m_) mkg_mp_destroy(m_);
    m_ = nullptr;
    check(mkg_mp_create(&mp_, &m_));
    const double R = std::max(3.0, std::min(mp_.nx, mp_.ny) / 6.0), W = mp_.width;
    const double cx = 0.5 * (mp_.nx - 1), cz = 0.5 * (mp_.nz - 1);
    std::vector<double> phi((size_t)mp_.nx * mp_.ny * mp_.nz);
    for (int x = 0; x < mp_.nx; x++)
        for (int y = 0; y < mp_.ny; y++)
            for (int z = 0; z < mp_.nz; z++) {
                const double dx = x - cx, dy = y + 0.5, dz = mp_.nz > 1 ? z - cz : 0.0;
                const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                phi[((size_t)x * mp_.ny + y) * mp_.nz + z] = 0.5 - 0.5 * std::tanh(2.0 * (r - R) / W);
            }
    check(mkg_mp_init(m_, phi.data(), nullptr, nullptr));
}

void mkFluidsMultiphase::advance(int mask) {
    if (!m_) throw mkt::mkLogicException("relaxation before initialize");
    if (mask & MKG_FLOW) check(mkg_mp_step(m_, 1));
    pending_ &= ~mask;
}

void mkFluidsMultiphase::write_vtkfile(int loops) {
    const int Mx = mp_.nx, My = mp_.ny, Mz = mp_.nz;
    const size_t N = (size_t)Mx * My * Mz;
    std::ostringstream fn;
    std::error_code ec;
    std::filesystem::create_directories(workspace_ + "vtk", ec);
    fn << workspace_ << "vtk/fluid_t" << loops << ".vtk";
    std::ofstream o(fn.str());
    if (!o) throw mkt::mkIOException("cannot write " + fn.str());
    o << "# vtk DataFile Version 3.0\nfluid_state\nASCII\nDATASET RECTILINEAR_GRID\n";
    o << "DIMENSIONS " << Mx << " " << My << " " << Mz << "\n";
    auto coords = [&](const char* nm, int m) {
        o << nm << " " << m << " float\n";
        for (int i = 0; i < m; i++) o << i + 0.5 << " ";
        o << "\n";
    };
    coords("X_COORDINATES", Mx);
    coords("Y_COORDINATES", My);
    coords("Z_COORDINATES", Mz);
    o << "POINT_DATA " << N << "\n" << std::setprecision(9);
    std::vector<double> phi(N), p(N), u(3 * N);
    check(mkg_mp_get(m_, MKG_MP_PHI, phi.data(), N));
    check(mkg_mp_get(m_, MKG_MP_PRESSURE, p.data(), N));
    check(mkg_mp_get(m_, MKG_MP_VEL, u.data(), 3 * N));
    auto scalar = [&](const char* name, const std::vector<double>& v) {
        o << "SCALARS " << name << " float 1\nLOOKUP_TABLE default\n";
        for (int z = 0; z < Mz; z++)
            for (int y = 0; y < My; y++)
                for (int x = 0; x < Mx; x++) o << v[((size_t)x * My + y) * Mz + z] << "\n";
    };
    scalar("order_parameter", phi);
    scalar("pressure", p);
    o << "VECTORS velocity_vector float\n";
    for (int z = 0; z < Mz; z++)
        for (int y = 0; y < My; y++)
            for (int x = 0; x < Mx; x++) {
                const size_t id = ((size_t)x * My + y) * Mz + z;
                o << u[3 * id] << " " << u[3 * id + 1] << " " << u[3 * id + 2] << "\n";
            }
}
