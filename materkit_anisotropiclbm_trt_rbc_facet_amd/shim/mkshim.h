// C++ host layer with the reference's class API (MaterKit/inc/mkFluidsLBM.h, mkFluidsMRT.h,
// mkFlowField.h, mkThermalField.h, mkSolutalField.h, mkPhaseField.h, mkSpace.h) over the
// C ABI in include/mkg.h, so that MaterKit-style drivers (MaterKit/MaterKit.cpp run_* loops)
// keep their call sequence.
//
// Semantics of the per-field calls:
//   X::propagate() / propagate_PBB()   mark field X for streaming (no device work)
//   X::relaxation*()                   one collide-stream of field X on the device; the other
//                                      fields are read at their current state for coupling
//                                      (mkg_step with that field's mask)
//   bdy*() / bdy_*()                   no device work: the face rules of the case file are
//                                      applied inside every collide-stream
//   step(n)                            n fused steps of every field (one kernel pair per step)
// Errors of the C ABI are thrown as mkt::mkIOException (case files) or
// mkt::mkLogicException (everything else), as the reference's exception classes
// (MaterKit/inc/mkRunTimeDiagnostics.h:71-117).
#pragma once
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "mkg.h"

namespace mkt {
struct mkIOException : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct mkLogicException : std::runtime_error {
    using std::runtime_error::runtime_error;
};
}  // namespace mkt

typedef double mkReal;  // the double-precision build of MaterKit/inc/mkGlobalDefs.h:175-180

// Three-vector of the reference (MaterKit/inc/mkVectors.h:123-221): the members the
// anisotropy queries and drivers use.
class mkVector3d {
public:
    mkVector3d() = default;
    mkVector3d(mkReal x, mkReal y, mkReal z) : v_{x, y, z} {}
    void set(mkReal x, mkReal y, mkReal z) {
        v_[0] = x;
        v_[1] = y;
        v_[2] = z;
    }
    mkReal& operator[](int i) { return v_[i]; }
    mkReal operator[](int i) const { return v_[i]; }
    mkReal* ptr() { return v_; }
    const mkReal* ptr() const { return v_; }
    mkReal x() const { return v_[0]; }
    mkReal y() const { return v_[1]; }
    mkReal z() const { return v_[2]; }
    mkReal operator*(const mkVector3d& r) const { return v_[0] * r[0] + v_[1] * r[1] + v_[2] * r[2]; }
    mkVector3d operator*(mkReal a) const { return {v_[0] * a, v_[1] * a, v_[2] * a}; }
    mkVector3d operator+(const mkVector3d& r) const { return {v_[0] + r[0], v_[1] + r[1], v_[2] + r[2]}; }
    mkVector3d operator-(const mkVector3d& r) const { return {v_[0] - r[0], v_[1] - r[1], v_[2] - r[2]}; }
    mkReal length_sq() const { return *this * *this; }
    mkReal length() const { return std::sqrt(length_sq()); }

private:
    mkReal v_[3] = {0, 0, 0};
};

// Grid of a case file (MaterKit/inc/mkSpace.h:29-73).
class mkSpace {
public:
    mkSpace() = default;
    explicit mkSpace(const std::string& xmlFileName) { reset(xmlFileName); }
    void reset(const std::string& xmlFileName);
    int Mx() const { return mx_; }
    int My() const { return my_; }
    int Mz() const { return mz_; }
    double step() const { return 1.0; }
    const std::string& xml() const { return xml_; }

private:
    int mx_ = 0, my_ = 0, mz_ = 0;
    std::string xml_;
};

// Populations of one node (MaterKit/inc/mkPDFs.h:22-123): up to 27 values, size() of them
// used (the lattice's Q).
class mkPDFs {
public:
    mkReal& operator[](int i) { return v_[i]; }
    mkReal operator[](int i) const { return v_[i]; }
    int size() const { return q_; }
    void resize(int q) { q_ = q; }
    mkReal density() const {
        mkReal s = 0;
        for (int i = 0; i < q_; i++) s += v_[i];
        return s;
    }

private:
    mkReal v_[27] = {};
    int q_ = 0;
};

// The reference's node record (MaterKit/inc/mkNode.h:134-198) as a host mirror of the device
// state: the pre-collision populations of every lattice at the current time level (f_hlp,
// g_hlp, h_hlp, x_hlp; canonical order of include/mkg.h's lattice tables, full values), the
// macroscopic fields (density, vel, phi, d_phi, temperature, X_solutes[0] = the conserved
// solute c), flag and grain_id.  The state itself lives on the device in structure-of-arrays
// form; mkFluidsLBM::fetch_nodes() materialises this mirror, store_nodes() writes the
// populations and flags back.
struct mkNode {
    mkPDFs f_hlp, g_hlp, h_hlp, x_hlp;
    mkVector3d vel;
    double density = 0, phi = 0, d_phi = 0;
    mkReal temperature = 0;
    std::vector<mkReal> X_solutes;
    int flag = 0;
    int grain_id = 0;
    int getGrainId() const { return grain_id; }
};

class mkFluidsLBM {
public:
    mkFluidsLBM() = default;
    explicit mkFluidsLBM(const mkSpace& space) : space_(space) {}
    virtual ~mkFluidsLBM();
    mkFluidsLBM(const mkFluidsLBM&) = delete;
    mkFluidsLBM& operator=(const mkFluidsLBM&) = delete;

    void reset(const std::string& xmlFileName) { space_.reset(xmlFileName); }
    void reset(const mkSpace& space) { space_ = space; }

    // parameters, device allocation and the equilibrium initial state of the case
    virtual void initialize(const std::string& xmlfname);
    // flags and seeds of the case (walls from the case's face codes)
    virtual void initialize_geometry();

    virtual void propagate() { pending_ |= MKG_FLOW; }
    virtual void propagate_PBB() { pending_ |= MKG_FLOW; }
    virtual void relaxation() { advance(MKG_FLOW); }
    virtual void bounceback() {}
    virtual void boundary() {}
    virtual void boundary_RBC() {}
    virtual void boundary_walls() {}
    virtual void bdyPeriod() {}
    virtual void bdyVelocity() {}
    virtual void bdyCouette() {}  // moving walls are the case file's face velocities
    // rotating / spiral walls are the case file's solver/wallRotation (+ the face velocities
    // for the axial part of a spiral), applied by the face rules in every step
    virtual void bdyRotation() {}
    virtual void bdySpiral() {}
    // parameters of a case file without creating the device state (initialize does both)
    virtual void read_parameters(const std::string& xmlfname);

    virtual double Thermal_NusseltRBC();
    virtual double Solutal_NusseltRBC();
    virtual void check_variables(int loops, const char nPlane);
    // legacy VTK frame (RECTILINEAR_GRID, cell centres x+0.5, x fastest), file
    // <workspace>vtk/fluid_t<loops>.vtk as read by the reference's SaveContourBatch.py:41
    virtual void write_vtkfile(int loops);
    // the mid-z plane as a legacy VTK frame, <workspace>vtk/fluid2D_t<loops>.vtk
    virtual void write_vtkfile_2D(int loops);
    // Tecplot ASCII (POINT) frames of the same fields: every cell, or the mid plane normal to
    // nPlane ('x', 'y', 'z'); <workspace>plt/fluid_<frame>.plt, fluid2D_<nPlane>_t<loops>.plt
    virtual void write_pltfile(int frame);
    virtual void write_pltfile_2D(int loops, char nPlane);

    // fused steps of every field
    void step(int nsteps = 1);
    // direct access for user code
    mkg_handle handle() const {
        fvm_flush();
        return h_;
    }
    const mkg_params& params() const { return p_; }
    std::vector<double> field(int id) const;
    void set_workspace(const std::string& dir) { workspace_ = dir; }
    // Node array of the reference (MaterKit/inc/mkFluidsLBM.h:95), materialised lazily:
    // null until fetch_nodes(), which copies the device state into a host mirror of
    // Mx*My*Mz records (index (x*My + y)*Mz + z) and points node at it; a later fetch_nodes()
    // refreshes it (device steps do not).  store_nodes() writes the mirror's populations and
    // flags back to the device (the macroscopic fields follow from the populations).
    mkNode* node = nullptr;
    mkNode* fetch_nodes();
    void store_nodes();

protected:
    virtual void advance(int mask);
    void check(int rc) const;
    // Finite-volume operators (mkPhaseField::evolution_anisotropy, mkSolutalField::evolution*):
    // an mkg_fvm handle per kind (solute without / with the anti-trapping current, or the
    // temperature as the potential), created at the first call from the case's parameters
    // (without grain rotations).  It takes phi and its potential (c or T) from the lattices
    // when it starts and keeps the newest values while only finite-volume calls follow; the
    // next lattice step or field read hands them back (the lattices restart at the
    // equilibrium of the current fields).  A phase call marks the next potential call to
    // advance phi with it.
    enum { kFvmSolute = 0, kFvmSoluteAT = 1, kFvmThermal = 2 };
    void fvm_evolve(int kind);
    void fvm_flush() const;
    mkg_fvm_handle fvm_[3] = {nullptr, nullptr, nullptr};
    mutable int fvm_live_ = -1;  // handle holding the newest phi and c; -1: the lattices
    bool fvm_phase_next_ = false;
    mkSpace space_;
    std::string xml_;
    std::string workspace_ = "./";
    mkg_handle h_ = nullptr;
    mkg_params p_{};
    std::vector<mkNode> nodes_;
    int32_t run_info_[4] = {0, 0, 0, 0};
    int pending_ = 0;
};

class mkFlowField : virtual public mkFluidsLBM {
public:
    mkFlowField() = default;
    explicit mkFlowField(const mkSpace& space) : mkFluidsLBM(space) {}
    void initialize_geometry_Poiseuille() { initialize_geometry(); }
};

class mkThermalField : virtual public mkFluidsLBM {
public:
    mkThermalField() = default;
    explicit mkThermalField(const mkSpace& space) : mkFluidsLBM(space) {}
    void propagate() override { pending_ |= MKG_THERMAL; }
    void relaxation() override { advance(MKG_THERMAL); }
    // the velocity coupling is read inside every thermal step
    void relaxation(const mkFluidsLBM& /*flow*/) { advance(MKG_THERMAL); }
    void update_velocity(const mkFluidsLBM& /*flow*/) {}
    void bdy_RBC() {}
    void bdy_noflux() {}
    void bdy_flowin() {}
    void bdy_constT() {}
};

class mkSolutalField : virtual public mkFluidsLBM {
public:
    mkSolutalField() = default;
    explicit mkSolutalField(const mkSpace& space) : mkFluidsLBM(space) {}
    void propagate() override { pending_ |= MKG_SOLUTE; }
    void propagate_PBB() override { pending_ |= MKG_SOLUTE; }
    void relaxation() override { advance(MKG_SOLUTE); }
    void relaxation_anisotropy() { advance(MKG_SOLUTE); }
    // finite-volume solute updates (MaterKit/inc/mkSolutalField.h:41-51): with the
    // anti-trapping current (ATC1 / ATC2) or without it (evolution; evolution_old, the
    // earlier entry of the same update); evolution_thermal takes the temperature as the
    // potential ("potential is temperature", :47) and needs the thermal lattice
    void evolution(const mkNode* /*nptr*/) { fvm_evolve(kFvmSolute); }
    void evolution_old(const mkNode* /*nptr*/) { fvm_evolve(kFvmSolute); }
    void evolution_ATC1(const mkNode* /*nptr*/) { fvm_evolve(kFvmSoluteAT); }
    void evolution_ATC2(const mkNode* /*nptr*/) { fvm_evolve(kFvmSoluteAT); }
    void evolution_thermal(const mkNode* /*nptr*/) { fvm_evolve(kFvmThermal); }
    void updateDCs(const mkNode* /*nptr*/) {}  // D(phi) is evaluated inside every solute update
    void bdy_constX() {}
    void bdy_RBC() {}
    void bdy_LBflowin() {}
};

class mkPhaseField : virtual public mkFluidsLBM {
public:
    mkPhaseField() = default;
    explicit mkPhaseField(const mkSpace& space) : mkFluidsLBM(space) {}
    void propagate() override { pending_ |= MKG_PHASE; }
    void relaxation() override { advance(MKG_PHASE); }
    void relaxation_anisotropy() { advance(MKG_PHASE); }
    void relaxation_anisotropy_thermal() { advance(MKG_PHASE); }
    void relaxation_anisotropy_thermal(const mkFluidsLBM& /*flow*/) { advance(MKG_PHASE); }
    void relaxation_anisotropy_solute(const int /*loops*/) { advance(MKG_PHASE); }
    // the anisotropy form is the case file's (<solver><anisotropyForm>); these two
    // check that it is the form their name asks for
    void relaxation_anisotropy_cbc() { aniso_form(MKG_ANISO_CBC, "cbc"); }
    void relaxation_anisotropy_hcp() { aniso_form(MKG_ANISO_HCP, "hcp"); }
    // finite-volume phase update (mkPhaseField.h:41,53): advances phi together with the
    // solute update that follows it
    void evolution_anisotropy() { fvm_phase_next_ = true; }
    void evolution(const mkNode* /*nptr*/) { fvm_phase_next_ = true; }
    // conservative Allen-Cahn interface (MaterKit/inc/mkPhaseField.h:39-40); the case file
    // must select <solver><phaseModel> allenCahn </phaseModel>
    void relaxation_AllenCahn() { allen_cahn(); }
    void relaxation_AllenCahn_growth_thermal() { allen_cahn(); }
    void reset_flags() {}  // flags are re-classified inside every phase step
    // Anisotropy of grain id (1-based, the node grain ids; 0 or no grains: unrotated) at the
    // interface normal (MaterKit/inc/mkPhaseField.h:61-64): A_s = a_s(n) and
    // N_v = R a_s (D - n (n.D)) of the case's strengths, in the form the name gives
    // (anisotropy: the cubic form).  x, y, z are not used: the grain id names the crystal.
    // False for a zero normal (A_s = 1, N_v = 0).
    bool anisotropy(const int& id, mkVector3d& N_v, mkReal& A_s, mkVector3d& normal,
                    const int& x, const int& y, const int& z) {
        return aniso_query(MKG_ANISO_CUBIC, id, N_v, A_s, normal);
    }
    bool anisotropy_facet(const int& id, mkVector3d& N_v, mkReal& A_s, mkVector3d& normal,
                          const int& x, const int& y, const int& z) {
        return aniso_query(MKG_ANISO_FACET, id, N_v, A_s, normal);
    }
    bool anisotropy_cbc(const int& id, mkVector3d& N_v, mkReal& A_s, mkVector3d& normal,
                        const int& x, const int& y, const int& z) {
        return aniso_query(MKG_ANISO_CBC, id, N_v, A_s, normal);
    }
    bool anisotropy_hcp(const int& id, mkVector3d& N_v, mkReal& A_s, mkVector3d& normal,
                        const int& x, const int& y, const int& z) {
        return aniso_query(MKG_ANISO_HCP, id, N_v, A_s, normal);
    }

private:
    bool aniso_query(int kind, int id, mkVector3d& N_v, mkReal& A_s, const mkVector3d& normal) {
        const double* R = (id > 0 && id <= p_.ngrains) ? p_.rot[id - 1] : nullptr;
        const int rc = mkg_aniso_eval(kind, p_.eps4, p_.aniso, R, normal.ptr(), &A_s, N_v.ptr());
        if (rc < 0) throw mkt::mkLogicException(mkg_last_error());
        return rc == 0;
    }
    void aniso_form(int kind, const char* name) {
        if (p_.aniso_kind != kind)
            throw mkt::mkLogicException(std::string("relaxation_anisotropy_") + name +
                                        ": the case file selects another anisotropy form");
        advance(MKG_PHASE);
    }
    void allen_cahn() {
        if (p_.pf_model != MKG_PF_ALLEN_CAHN)
            throw mkt::mkLogicException(
                "relaxation_AllenCahn: the case file selects the Karma-Rappel phase model "
                "(<solver><phaseModel> allenCahn </phaseModel> selects Allen-Cahn)");
        advance(MKG_PHASE);
    }
};

class mkFluidsMRT : virtual public mkFluidsLBM,
                    virtual public mkFlowField,
                    virtual public mkThermalField,
                    virtual public mkSolutalField,
                    virtual public mkPhaseField {
public:
    mkFluidsMRT() = default;
    explicit mkFluidsMRT(const mkSpace& space) : mkFluidsLBM(space) {}
    void propagate() override { pending_ |= MKG_FLOW; }
    void propagate_PBB() override { pending_ |= MKG_FLOW; }
    void relaxation() override { advance(MKG_FLOW); }
    // the node-array form (MaterKit/inc/mkFluidsMRT.h:46): the node array lives on the device
    void relaxation(const mkNode* /*nptr*/) { advance(MKG_FLOW); }
    void initialize(const std::string& xmlfname) override { mkFluidsLBM::initialize(xmlfname); }
    void initialize_geometry() override { mkFluidsLBM::initialize_geometry(); }
};

// Multi-phase field of the user-extension class (MaterKit/inc/mkFluidsUser.h;
// propagate / relaxation of MaterKit/src/mkFluidsUser.cpp:68-299): one order parameter per
// seed of the case file (at most MKG_MPF_MAX), each on its own D3Q19 lattice, plus the
// case's D3Q7 thermal lattice, on an mkg_mpf handle.  relaxation() advances every lattice by
// one step; the thermal calls of run_multiPhaseField (mkThermalField::propagate / relaxation)
// are then no-ops, because the thermal lattice steps with the order parameters.
class mkFluidsUser : public mkFluidsMRT {
public:
    mkFluidsUser() = default;
    explicit mkFluidsUser(const mkSpace& space) : mkFluidsLBM(space), mkFluidsMRT(space) {}
    ~mkFluidsUser() override;
    void initialize(const std::string& xmlfname) override;
    void initialize_geometry() override;
    void propagate() override { pending_ |= MKG_PHASE; }
    void relaxation() override { advance(MKG_PHASE); }
    void write_vtkfile(int loops) override;
    mkg_mpf_handle mpf() const { return m_; }
    const mkg_mpf_params& mpf_params() const { return mp_; }
    // phi_j (n x N) and the combined phi (N) at the last moment pass
    std::vector<double> order_parameters() const;

protected:
    void advance(int mask) override;

private:
    mkg_mpf_handle m_ = nullptr;
    mkg_mpf_params mp_{};
};

// Anisotropic multi-grain phase field (MaterKit/inc/mkMultiphaseField.h:23-58) on an mkg_mpf
// handle with Karma-Rappel grains: one order parameter per seed of the case file (at most
// MKG_MPF_MAX), each with the case's anisotropy form and strengths (<phasechange>
// <anisotropy>, <solver><anisotropyForm>) and the seed's own orientation (the ZYX Euler
// angles of <seed><angle>, one triple per seed; a seed without one is unrotated), driven by
// the case's thermal
// lattice (relaxation_anisotropy_thermal, relaxation_thermal) or solute lattice
// (relaxation_solute, for a case with <solver><solute> and no thermal lattice or
// <solver><coupling> solute).  Each relaxation call advances every lattice by one step; the
// driving field is the case's, and a call naming the other one throws.  The finite-volume
// evolution_anisotropy and the Allen-Cahn relaxation of this class are not provided (they
// throw): the Allen-Cahn grains are mkFluidsUser's.
class mkMultiphaseField : public mkFluidsMRT {
public:
    mkMultiphaseField() = default;
    explicit mkMultiphaseField(const mkSpace& space) : mkFluidsLBM(space), mkFluidsMRT(space) {}
    ~mkMultiphaseField() override;
    void initialize(const std::string& xmlfname) override;
    void initialize_geometry() override;
    void propagate() override { pending_ |= MKG_PHASE; }
    void relaxation() override { advance(MKG_PHASE); }
    void relaxation_anisotropy_thermal() { driven(1, "relaxation_anisotropy_thermal"); }
    void relaxation_thermal() { driven(1, "relaxation_thermal"); }
    void relaxation_solute() { driven(2, "relaxation_solute"); }
    void relaxation_AllenCahn();
    void evolution_anisotropy();
    void evolution(const mkNode* /*nptr*/) { evolution_anisotropy(); }
    // Anisotropy of order parameter id - 1 (MaterKit/inc/mkMultiphaseField.h:31-34): its
    // rotation, the case's strengths, the form the name gives.  False for a zero normal.
    bool anisotropy(const int& id, mkVector3d& N_v, mkReal& A_s, mkVector3d& normal,
                    const int& x, const int& y, const int& z) {
        return grain_query(MKG_ANISO_CUBIC, id, N_v, A_s, normal);
    }
    bool anisotropy_facet(const int& id, mkVector3d& N_v, mkReal& A_s, mkVector3d& normal,
                          const int& x, const int& y, const int& z) {
        return grain_query(MKG_ANISO_FACET, id, N_v, A_s, normal);
    }
    bool anisotropy_cbc(const int& id, mkVector3d& N_v, mkReal& A_s, mkVector3d& normal,
                        const int& x, const int& y, const int& z) {
        return grain_query(MKG_ANISO_CBC, id, N_v, A_s, normal);
    }
    bool anisotropy_hcp(const int& id, mkVector3d& N_v, mkReal& A_s, mkVector3d& normal,
                        const int& x, const int& y, const int& z) {
        return grain_query(MKG_ANISO_HCP, id, N_v, A_s, normal);
    }
    void write_vtkfile(int loops) override;
    mkg_mpf_handle mpf() const { return m_; }
    const mkg_mpf_params& mpf_params() const { return mp_; }
    // phi_j (n x N) and the combined phi (N) at the last moment pass
    std::vector<double> order_parameters() const;

protected:
    void advance(int mask) override;
    int seed_number = 0;   // seeds of the case file
    int phase_number = 0;  // order parameters (one per seed)
    std::vector<double> vec_angle, vec_center;

private:
    void driven(int coupling, const char* name);
    bool grain_query(int kind, int id, mkVector3d& N_v, mkReal& A_s, const mkVector3d& normal);
    mkg_mpf_handle m_ = nullptr;
    mkg_mpf_params mp_{};
};

// Phase-field two-phase flow (MaterKit/inc/mkFluidsMultiphase.h:20-99) on an mkg_mp handle:
// parameters from the case file's <multiphases> block (mkg_mp_params_from_xml).  The
// geometry is a layered channel (the dense fluid in the lower half along y), or with
// <solver><twoPhaseInit> drops </twoPhaseInit> one drop of the dense fluid per seed.  The
// reference ships this class's bodies only in compiled form, so the start geometry is this
// library's own choice.
// relaxation() advances both lattices by one step.
class mkFluidsMultiphase : public mkFluidsMRT {
public:
    mkFluidsMultiphase() = default;
    explicit mkFluidsMultiphase(const mkSpace& space) : mkFluidsLBM(space), mkFluidsMRT(space) {}
    ~mkFluidsMultiphase() override;
    void initialize(const std::string& xmlfname) override;
    void initialize_geometry() override;
    // The case's parameters with wetting walls (run_flow_multiphase_wetting,
    // MaterKit/MaterKit.cpp:333-375): the y faces become walls when the case has none, every
    // wall carries multiphases/contactAngle, and a half-disc (half-sphere in 3D) drop of the
    // dense fluid of radius min(nx, ny) / 6 sits on the lower y wall at the centre of x (and z).
    void initialize_wetting(const std::string& xmlfname);
    void propagate() override { pending_ |= MKG_FLOW; }
    void relaxation() override { advance(MKG_FLOW); }
    void relaxation_phi() {}  // the Cahn-Hilliard lattice steps with the flow
    // The solidifying two-phase coupling of run_flow_multiphase_dendrite
    // (MaterKit/inc/mkFluidsMultiphase.h:37) is not provided: it throws.
    void update_geometry(mkFluidsMRT& dendrite);
    void write_vtkfile(int loops) override;
    mkg_mp_handle mp() const { return m_; }
    const mkg_mp_params& mp_params() const { return mp_; }

protected:
    void advance(int mask) override;

private:
    mkg_mp_handle m_ = nullptr;
    mkg_mp_params mp_{};
};

// readParameters of the reference (MaterKit/inc/mkFunction.h:62): loops and frames.
void readParameters(int& maxLoops, int& maxFrame, const std::string& xmlFileName);

// Driver utilities of the reference that its run_* loops use: the wall-clock timer
// (MaterKit/inc/mkCPUTimer.h:36-55, the global `timer` of AiLBFlow.h:83), which prints its
// label and the elapsed seconds when deleted; the case-file name (mkt::paraXmlFileName,
// mkFunction.h:132); the output folder prefix (workspace, mkGlobalDefs.h:86: here the current
// directory); and main's set-up calls (mkFunction.h:58-59): make_datafolder creates the vtk
// folder under the workspace, make_numthreads returns 1 (the device runs the steps; no host
// threads).
class mkCPUTimer {
public:
    mkCPUTimer() : mkCPUTimer("timer") {}
    explicit mkCPUTimer(const std::string label);
    ~mkCPUTimer();
    double GetTime(int flag);  // seconds since construction (flag unused)
    double GetSystemTime();    // seconds since the epoch

private:
    std::string label_;
    std::chrono::steady_clock::time_point t0_;
};
extern mkCPUTimer* timer;
namespace mkt {
extern std::string paraXmlFileName;
}
extern const std::string workspace;
extern "C" {
int make_datafolder();
int make_numthreads();
}
