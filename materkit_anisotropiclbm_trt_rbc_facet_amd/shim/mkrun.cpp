// MaterKit-style driver over the reference-API layer: the run_* loops of
// MaterKit/MaterKit.cpp (run_flow_Poiseuille :244-288, run_flow_solutal_dendrite :399-465,
// run_flow_thermal_dendrite :467-527, run_solutal_dendrite :633-687, run_solutal_only :689-741,
// run_thermal_dendrite :743-788) with their per-step call
// order, plus a fused loop.  The case file and the loop are chosen on the command line
// (the reference hard-codes choice = 8, MaterKit.cpp:50).
//
//   mkrun <case.xml> <loop> [maxLoops] [maxFrame] [vtk | plt]
//   loop: poiseuille | thermal_dendrite | flow_thermal_dendrite | flow_solutal_dendrite |
//         solutal_dendrite | solutal_only | thermal_fvm | rbc | allen_cahn |
//         multi_phase_field | grains | multiphase | wetting | fused
// (allen_cahn: run_flow_AllenCahn, MaterKit.cpp:186-242; multi_phase_field:
// run_multiPhaseField, MaterKit.cpp:529-576, on mkFluidsUser; grains: the same loop on the
// anisotropic multi-grain class mkMultiphaseField with relaxation_anisotropy_thermal or
// relaxation_solute, as the case drives the grains; multiphase: run_flow_multiphase,
// MaterKit.cpp:290-331, and wetting: run_flow_multiphase_wetting, :333-375, on
// mkFluidsMultiphase; thermal_fvm: evolution_anisotropy and evolution_thermal, the
// finite-volume phase field with the temperature as its potential)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mkshim.h"

static void frame_info(mkFluidsMRT& flow, int frame, int loops) {
    std::printf("%05d\t%18d", frame, loops);
    const mkg_params& p = flow.params();
    if (p.phase_q) {
        double v[4];
        mkg_reduce(flow.handle(), MKG_R_SOLID, v);
        std::printf("\tsolid_fraction %.8e", v[0]);
        mkg_reduce(flow.handle(), MKG_R_TIP, v);
        std::printf("\ttip_x %.3f", v[0]);
    }
    if (p.thermal && p.flow_q && p.T_face[4] != p.T_face[5])
        std::printf("\tNu %.8f", flow.Thermal_NusseltRBC());
    std::printf("\n");
}

// run_multiPhaseField (MaterKit/MaterKit.cpp:529-576) on the multi-phase user class
static int run_multi_phase_field(const std::string& xml, int maxLoops, int maxFrame, bool vtk) {
    mkSpace space(xml);
    mkFluidsUser flow(space);
    flow.initialize(xml);
    flow.initialize_geometry();
    std::printf("%20s= %d \n%20s= %d \n\nframe \t\t\t loops\n", "maxLoops ", maxLoops,
                "maxFrame ", maxFrame);
    const int every = std::max(1, maxLoops / maxFrame);
    const mkg_mpf_params& m = flow.mpf_params();
    std::vector<double> sums(m.n + 1);
    int frame = 0;
    auto t0 = std::chrono::steady_clock::now();
    for (int loops = 0; loops <= maxLoops; loops++) {
        if (loops % every == 0) {
            mkg_mpf_sums(flow.mpf(), sums.data());
            std::printf("%05d\t%18d", frame, loops);
            for (int j = 0; j < m.n; j++) std::printf("\tsum_phi_%d %.8e", j, sums[j]);
            std::printf("\n");
            if (vtk) flow.write_vtkfile(loops);
            frame++;
        }
        if (loops == maxLoops) break;
        flow.propagate();
        flow.relaxation();
        flow.mkThermalField::propagate();
        flow.mkThermalField::relaxation();
    }
    double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("wall %.3f s, %.1f MLUPS (including frame output)\n", s,
                (double)space.Mx() * space.My() * space.Mz() * maxLoops / s / 1e6);
    return 0;
}

// The multi-phase loop on the anisotropic multi-grain class: per frame the solid cells
// (phi_j > 0) of every grain
static int run_grains(const std::string& xml, int maxLoops, int maxFrame, bool vtk) {
    mkSpace space(xml);
    mkMultiphaseField flow(space);
    flow.initialize(xml);
    flow.initialize_geometry();
    std::printf("%20s= %d \n%20s= %d \n\nframe \t\t\t loops\n", "maxLoops ", maxLoops,
                "maxFrame ", maxFrame);
    const int every = std::max(1, maxLoops / maxFrame);
    const mkg_mpf_params& m = flow.mpf_params();
    const size_t N = (size_t)m.nx * m.ny * m.nz;
    int frame = 0;
    auto t0 = std::chrono::steady_clock::now();
    for (int loops = 0; loops <= maxLoops; loops++) {
        if (loops % every == 0) {
            const std::vector<double> ph = flow.order_parameters();
            std::printf("%05d\t%18d", frame, loops);
            for (int j = 0; j < m.n; j++) {
                size_t solid = 0;
                for (size_t k = 0; k < N; k++) solid += ph[(size_t)j * N + k] > 0.0;
                std::printf("\tsolid_%d %zu", j, solid);
            }
            std::printf("\n");
            if (vtk) flow.write_vtkfile(loops);
            frame++;
        }
        if (loops == maxLoops) break;
        flow.propagate();
        if (m.coupling == 2)
            flow.relaxation_solute();
        else
            flow.relaxation_anisotropy_thermal();
    }
    double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("wall %.3f s, %.1f MLUPS (including frame output)\n", s,
                (double)N * maxLoops / s / 1e6);
    return 0;
}

// run_flow_multiphase (MaterKit/MaterKit.cpp:290-331) on the two-phase class
static int run_flow_multiphase(const std::string& xml, int maxLoops, int maxFrame, bool vtk,
                               bool wetting = false) {
    mkSpace space(xml);
    mkFluidsMultiphase flow(space);
    if (wetting) {
        flow.initialize_wetting(xml);
    } else {
        flow.initialize(xml);
        flow.initialize_geometry();
    }
    std::printf("%20s= %d \n%20s= %d \n\nframe \t\t\t loops\n", "maxLoops ", maxLoops,
                "maxFrame ", maxFrame);
    const int every = std::max(1, maxLoops / maxFrame);
    const mkg_mp_params& m = flow.mp_params();
    const size_t N = (size_t)m.nx * m.ny * m.nz;
    std::vector<double> u(3 * N);
    int frame = 0;
    auto t0 = std::chrono::steady_clock::now();
    for (int loops = 0; loops <= maxLoops; loops++) {
        if (loops % every == 0) {
            double s = 0, umax = 0;
            mkg_mp_sums(flow.mp(), &s);
            mkg_mp_get(flow.mp(), MKG_MP_VEL, u.data(), u.size());
            for (size_t k = 0; k < N; k++) umax = std::max(umax, std::fabs(u[3 * k]));
            std::printf("%05d\t%18d\tsum_phi %.10e\tumax_x %.8e\n", frame, loops, s, umax);
            if (vtk) flow.write_vtkfile(loops);
            frame++;
        }
        if (loops == maxLoops) break;
        flow.propagate();
        flow.relaxation();
    }
    double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("wall %.3f s, %.1f MLUPS (including frame output)\n", s,
                (double)N * maxLoops / s / 1e6);
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <case.xml> <loop> [maxLoops] [maxFrame] [vtk]\n", argv[0]);
        return 2;
    }
    const std::string xml = argv[1], loop = argv[2];
    int maxLoops = 0, maxFrame = 0;
    try {
        readParameters(maxLoops, maxFrame, xml);
        if (argc > 3) maxLoops = std::atoi(argv[3]);
        if (argc > 4) maxFrame = std::atoi(argv[4]);
        const bool vtk = argc > 5 && std::strcmp(argv[5], "vtk") == 0;
        // "plt": Tecplot frames (every cell) and the mid-z plane as VTK and Tecplot
        const bool plt = argc > 5 && std::strcmp(argv[5], "plt") == 0;
        if (maxFrame < 1) maxFrame = 1;
        if (loop == "multi_phase_field") return run_multi_phase_field(xml, maxLoops, maxFrame, vtk);
        if (loop == "grains") return run_grains(xml, maxLoops, maxFrame, vtk);
        if (loop == "multiphase") return run_flow_multiphase(xml, maxLoops, maxFrame, vtk);
        if (loop == "wetting") return run_flow_multiphase(xml, maxLoops, maxFrame, vtk, true);
        mkSpace space(xml);
        mkFluidsMRT flow(space);
        flow.mkFluidsMRT::initialize(xml);
        if (loop == "poiseuille")
            flow.mkFluidsMRT::initialize_geometry_Poiseuille();
        else
            flow.mkFluidsMRT::initialize_geometry();
        std::printf("%20s= %d \n%20s= %d \n\nframe \t\t\t loops\n", "maxLoops ", maxLoops,
                    "maxFrame ", maxFrame);
        const int every = std::max(1, maxLoops / maxFrame);
        int frame = 0;
        auto t0 = std::chrono::steady_clock::now();
        for (int loops = 0; loops <= maxLoops; loops++) {
            if (loops % every == 0) {
                frame_info(flow, frame, loops);
                if (vtk) flow.write_vtkfile(loops);
                if (plt) {
                    flow.write_pltfile(frame);
                    flow.write_pltfile_2D(loops, 'z');
                    flow.write_vtkfile_2D(loops);
                }
                frame++;
            }
            if (loops == maxLoops) break;
            if (loop == "poiseuille") {
                flow.mkFluidsLBM::propagate();
                flow.mkFluidsLBM::relaxation();
            } else if (loop == "thermal_dendrite") {
                flow.mkPhaseField::propagate();
                flow.mkPhaseField::relaxation_anisotropy_thermal();
                flow.mkThermalField::propagate();
                flow.mkThermalField::relaxation();
            } else if (loop == "flow_thermal_dendrite") {
                flow.mkFluidsMRT::propagate_PBB();
                flow.mkFluidsMRT::relaxation();
                flow.mkFluidsMRT::bdyVelocity();
                flow.mkPhaseField::propagate();
                flow.mkPhaseField::relaxation_anisotropy_thermal();
                flow.mkThermalField::propagate();
                flow.mkThermalField::relaxation();
                flow.mkThermalField::bdy_flowin();
            } else if (loop == "flow_solutal_dendrite") {
                flow.mkFluidsMRT::propagate_PBB();
                flow.mkFluidsMRT::relaxation();
                flow.bdyVelocity();
                flow.mkPhaseField::propagate();
                flow.mkPhaseField::relaxation_anisotropy_solute(loops);
                flow.mkSolutalField::propagate();
                flow.mkSolutalField::relaxation();
                flow.mkSolutalField::bdy_LBflowin();
            } else if (loop == "solutal_dendrite") {
                flow.mkPhaseField::propagate();
                flow.mkPhaseField::relaxation_anisotropy_solute(loops);
                flow.mkSolutalField::propagate();
                flow.mkSolutalField::relaxation();
            } else if (loop == "solutal_only") {
                flow.mkSolutalField::evolution_ATC2(flow.node);
            } else if (loop == "thermal_fvm") {
                // finite-volume dendrite: the phase field and the temperature as the potential
                flow.mkPhaseField::evolution_anisotropy();
                flow.mkSolutalField::evolution_thermal(flow.node);
            } else if (loop == "rbc") {
                flow.mkFluidsMRT::propagate();
                flow.mkFluidsMRT::relaxation();
                flow.mkThermalField::propagate();
                flow.mkThermalField::relaxation();
                flow.mkThermalField::bdy_RBC();
            } else if (loop == "allen_cahn") {
                flow.mkFluidsMRT::propagate();
                flow.mkFluidsMRT::relaxation();
                flow.mkPhaseField::propagate();
                flow.mkPhaseField::relaxation_AllenCahn();
            } else if (loop == "fused") {
                flow.step(1);
            } else {
                std::fprintf(stderr, "unknown loop %s\n", loop.c_str());
                return 2;
            }
        }
        double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const double cells = (double)space.Mx() * space.My() * space.Mz();
        std::printf("wall %.3f s, %.1f MLUPS (including frame output)\n", s,
                    cells * maxLoops / s / 1e6);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
