/*
 * mk_oracle.h -- FP64 CPU oracle for the MaterKit lattice-Boltzmann / phase-field step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the step, written for
 * readability: structure-of-arrays doubles, two-lattice pull streaming, textbook per-cell
 * formulas.  Only the test-suite, the smoke check and the CPU-baseline leg of the benchmark
 * may load it; the product library never links it.
 *
 * The reference's solver bodies exist only inside a closed Windows DLL
 * (MaterKit/AiLBFlow.dll), so the formulas follow the reference headers and the visible
 * user-extension loops (MaterKit/src/mkFluidsUser.cpp) plus the published methods named at
 * each function in mk_oracle.c.  The parameter block has the same layout as the product's
 * mkg_params so one Python description drives both.
 */
#ifndef MK_ORACLE_H
#define MK_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#define ORC_MAX_GRAINS 64

typedef struct orc_params {
    int32_t nx, ny, nz;
    int32_t flow_q, phase_q, thermal, solute, precision, flow_op, pf_coupling, ngrains,
        has_kappa;
    int32_t bc_f[6], bc_g[6], bc_h[6], bc_c[6];
    int32_t flow_incomp;
    double s_nu, s_bulk, s_odd, s_even;
    double rho0;
    double force[3];
    double buoy[3];
    double T_ref;
    double u_face[6][3];
    double alpha_l, alpha_s, latent_T;
    double T_face[6];
    double D_l, D_s, k_part, latent_C, q_min;
    double at_coef;
    double C_face[6];
    double pf_dt, pf_wdx2, eps4, lambda, phi_s;
    int32_t aniso_kind, pf_model;
    double aniso[2];
    double ref_f, ref_g, ref_h, ref_c;
    double rot[ORC_MAX_GRAINS][9];
    double wall_omega[3];
    double wall_center[3];
} orc_params;

/* Collision operator selector for the flow lattice (the oracle keeps three separate
 * implementations so each can be checked against the others' special cases). */
#define ORC_OP_BGK 0
#define ORC_OP_TRT 1
#define ORC_OP_MRT 2

typedef struct orc_state {
    double* f;          /* flow_q  x N, pre-collision populations (full values) */
    double* g;          /* phase_q x N */
    double* h;          /* 7 x N */
    double* c;          /* 7 x N */
    double* phi;        /* N: phase field of the last phase moment pass */
    double* phi_prev;   /* N: the pass before */
    int8_t* flag;       /* N */
    int16_t* grain;     /* N */
    const double* kappa;/* N or NULL */
} orc_state;

size_t orc_params_size(void);

/* One or more steps advancing the lattices in mask (1 flow, 2 phase, 4 thermal, 8 solute);
 * op selects the flow collision implementation (ORC_OP_*). */
int orc_step(const orc_params* p, orc_state* s, int mask, int op, int nsteps);

/* Macroscopic fields at the current time level (any pointer may be NULL). */
void orc_moments(const orc_params* p, const orc_state* s, double* rho, double* vel3,
                 double* T, double* C, double* phi_now);

/* Equilibrium initial state from macroscopic fields (NULL = uniform defaults). */
void orc_init_equilibrium(const orc_params* p, orc_state* s, const double* rho,
                          const double* vel3, const double* T, const double* C,
                          const double* phi);

/* Seed geometry: phi profile, grain ids (nearest seed surface), flags. */
void orc_init_seeds(const orc_params* p, orc_state* s, int n, const double* centers,
                    const double* radii);

/* Single-cell helpers exposed for unit tests. */
void orc_equilibrium(int q, double rho, const double* u, double* feq);
void orc_equilibrium_he(int q, double rho, double rho0, const double* u, double* feq);
void orc_collide_cell(const orc_params* p, int op, const double* f, const double* F, double* out,
                      double ns);
void orc_phase_source(const orc_params* p, const double* phi27, double upf, int grain,
                      double* S_out, double* a_out);

double orc_face_flux(const orc_params* p, const double* R, const double* g, int axis);
double orc_aniso(const orc_params* p, const double* n, double* D);

/* Reference-structured step (mk_aos.c): the same step as orc_step(mask 15) on AoS node
 * records of the reference's layout, swept operator by operator; for periodic flow (D3Q19) +
 * phase + thermal cases without walls.  *seconds = time of the step loop alone.  Returns -1
 * outside that scope. */
int orc_aos_step(const orc_params* p, orc_state* s, int op, int nsteps, double* seconds);

/* Multi-phase field (mkFluidsUser's phase_number order parameters): n order parameters on
 * D3Q19 with pairwise interaction -- conservative Allen-Cahn (the template) or Karma-Rappel
 * grains, each with its own anisotropy and orientation (mkMultiphaseField) -- and optional
 * D3Q7 thermal and solute lattices.  Same layout as the product's mkg_mpf_params. */
#define ORC_MPF_MAX 4
typedef struct orc_mpf_params {
    int32_t nx, ny, nz, n;
    int32_t precision, thermal;
    int32_t bc[6], bc_h[6];
    double dt, wdx;                 /* pf_dt and W0/dx */
    double u[ORC_MPF_MAX][3];       /* advection velocity of each order parameter */
    double kappa[ORC_MPF_MAX];      /* interaction strength of each */
    double growth, lambda, eps;     /* growth-term coefficient, coupling, normal regulariser */
    double alpha, latent_T, T_ref;  /* thermal lattice */
    double T_face[6];
    double phi_face[6];             /* Dirichlet value of the phase lattices */
    int32_t model;                  /* 0 template (Allen-Cahn), 1 Karma-Rappel grains */
    int32_t coupling;               /* field driving the grains: 0 none, 1 T, 2 U */
    int32_t aniso_kind[ORC_MPF_MAX];
    double eps4[ORC_MPF_MAX];
    double aniso[ORC_MPF_MAX][2];
    double rot[ORC_MPF_MAX][9];     /* lab <- crystal rotation of each grain, row-major */
    int32_t solute;                 /* D3Q7 solute lattice driven by the combined phi */
    int32_t bc_c[6];
    double D_l, D_s, k_part, q_min, latent_C, at_coef;
    double C_face[6];
} orc_mpf_params;

/* g: n x 19 x N populations; phi, phi_prev: (n + 1) x N (phi_j, then the combined phi);
 * h, c: 7 x N or NULL; T, C: the initial temperature and solute c (NULL: T_ref, 0);
 * flag: N or NULL (no walls). */
void orc_mpf_init(const orc_mpf_params* m, const double* phij, const double* T, const double* C,
                  double* g, double* phi, double* phi_prev, double* h, double* c);
int orc_mpf_step(const orc_mpf_params* m, double* g, double* phi, double* phi_prev, double* h,
                 double* c, int8_t* flag, int nsteps);

/* Explicit finite-volume phase field (evolution_anisotropy) and solute with anti-trapping
 * current (evolution_ATC1/2): phi and c (= A(phi) U) are N doubles each, advanced in place;
 * evolve_phase 0 holds phi fixed (run_solutal_only).  Uses the phase and solute parameters of
 * p (pf_dt, pf_wdx2, eps4, lambda, aniso_*, D_l, D_s, k_part, at_coef, q_min, bc_g, bc_c,
 * C_face as U values). */
int orc_fvm_step(const orc_params* p, double* phi, double* c, int nsteps, int evolve_phase);

/* Finite-volume phase field driven by the temperature, and the heat equation with the latent
 * source (evolution_thermal): phi and T are N doubles each, advanced in place.  Uses the
 * phase parameters as orc_fvm_step does, with u = T, and alpha_l, alpha_s, latent_T, bc_h,
 * T_face. */
int orc_fvm_thermal_step(const orc_params* p, double* phi, double* T, int nsteps,
                         int evolve_phase);

/* Two-phase flow (mkFluidsMultiphase; Liang et al. 2014 phase-field model): Cahn-Hilliard
 * lattice h and velocity-pressure flow lattice f, both D3Q19.  Same layout as the product's
 * mkg_mp_params. */
typedef struct orc_mp_params {
    int32_t nx, ny, nz, precision;
    int32_t bc[6];                 /* periodic (0) or bounce (1): no-slip wall / zero flux */
    double rho_l, rho_g;           /* densities of phi = 1 and phi = 0 */
    double nu_l, nu_g;             /* kinematic viscosities */
    double sigma, width;           /* surface tension, interface width W */
    double mobility, tau_h;        /* M and the Cahn-Hilliard relaxation time */
    double gravity[3], force[3];   /* body acceleration and body-force density */
    int32_t wet_faces;             /* bit f: bounce face f is a wetting wall */
    int32_t visc_model;            /* 0 linear nu, 1 linear rho nu, 2 harmonic rho nu */
    double contact_angle;          /* degrees, on the wetting faces */
} orc_mp_params;

/* f, h: 19 x N; phiu: 3 x N (phi u of the last step) */
void orc_mp_init(const orc_mp_params* m, const double* phi0, const double* u0, const double* p0,
                 double* f, double* h, double* phiu);
int orc_mp_step(const orc_mp_params* m, double* f, double* h, double* phiu, int nsteps);
/* phi, mu, p (N each) and u (3 x N, cell-interleaved) of the current state */
void orc_mp_macro(const orc_mp_params* m, const double* f, const double* h, double* phi,
                  double* mu, double* p, double* u3);

/* Discrete adjoint of nsteps flow-only steps (generalised moment collision, Guo forcing;
 * faces periodic, bounce or outflow; wall cells) from the pre-collision populations f0
 * (q x N): L = 1/2 sum over non-wall cells |u_N - u_target|^2; grad[7] = dL/dF (3),
 * dL/d(s_nu, s_bulk, s_odd, s_even).  Returns -1 for an unsupported configuration. */
int orc_flow_adjoint(const orc_params* p, const int8_t* flag, const double* f0, int nsteps,
                     const double* u_target, double* loss, double* grad);

/* Lattice tables (for tests). */
int orc_lattice(int q, int* e /* q*3 */, double* w, int* inv, double* cs2);

/* splitmix64 helpers shared with the golden-vector fixtures. */
uint64_t orc_splitmix64(uint64_t* state);
double orc_uniform(uint64_t* state);
double orc_cell_uniform(uint64_t seed, uint64_t cell);

#endif
