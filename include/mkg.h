/*
 * mkg.h -- C ABI of the MI355X (gfx950) lattice-Boltzmann / phase-field time step.
 *
 * This is the drop-in boundary for MaterKit's per-step solver calls.  Every entry point is
 * plain C (extern "C"), takes plain pointers and sizes, returns an int status (0 = ok) and
 * never throws; the text of the last error is available from mkg_last_error().
 *
 * The handle owns all device memory (structure-of-arrays populations, phase field, flags,
 * per-cell conductivity).  Calls are issued on one HIP stream per handle; get/reduce calls
 * synchronise that stream.  A handle is not thread-safe; separate handles are independent.
 *
 * Reference interfaces replaced (MaterKit, all bodies of which live in a closed DLL):
 *   mkg_params_from_xml   <- mkFluidsLBM::read_parameters / mkFluidsMRT::initialize(xml)
 *                            (MaterKit/inc/mkFluidsLBM.h:41-43, mkFluidsMRT.h:39),
 *                            readParameters (MaterKit/inc/mkFunction.h:62)
 *   mkg_create            <- mkFluidsMRT(const mkSpace&) + node[] allocation
 *                            (MaterKit/inc/mkFluidsMRT.h:37, mkFluidsLBM.h:95)
 *   mkg_init_seeds,       <- initialize_geometry / initialize_geometry_Poiseuille /
 *   mkg_init_equilibrium,    initialize_seeds (mkFluidsMRT.h:40, mkFlowField.h:33,
 *   mkg_set_field(FLAG)      mkMultiphaseField.h:52, MaterKit/src/mkFluidsUser.cpp:44-66):
 *                            seed spheres, equilibrium start, wall cells
 *   mkg_step (mask)       <- per-step call sequence propagate()/propagate_PBB()/relaxation()
 *                            of every field (MaterKit/MaterKit.cpp:277-280, 500-518,
 *                            775-781; mkFluidsLBM.h:47-49; mkPhaseField.h:41-48;
 *                            mkThermalField.h:33-35; mkSolutalField.h:35-37)
 *   face boundary codes   <- bdyVelocity / bdy_RBC / bdy_flowin / bdy_constT / bdy_LBflowin
 *                            (mkFluidsLBM.h:58-65, mkThermalField.h:37-40,
 *                            mkSolutalField.h:51-53), applied inside the step
 *   mkg_get_field         <- node[index].{density, vel, phi, temperature, flag, ...}
 *                            (MaterKit/inc/mkNode.h:134-198)
 *   mkg_reduce            <- Thermal_NusseltRBC / Solutal_NusseltRBC / check_variables
 *                            (mkFluidsLBM.h:55,67-68)
 *   mkg_checkpoint_*      <- (no counterpart; the reference writes VTK/Tecplot frames only)
 *   mkg_adjoint_*         <- (no counterpart; discrete adjoint of the step)
 */
#ifndef MKG_H
#define MKG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MKG_ABI_VERSION 5

/* ---- node flags: identical codes to MaterKit/inc/mkGlobalDefs.h:140-145 ---- */
#define MKG_IS_OUTLET (-2)
#define MKG_IS_INLET (-1)
#define MKG_IS_FLUID 0
#define MKG_IS_INTER 1
#define MKG_IS_SOLID 2
#define MKG_IS_WALLS 3

/* ---- lattices / operators / precision ---- */
#define MKG_OP_BGK 0
#define MKG_OP_TRT 1
#define MKG_OP_MRT 2
#define MKG_PREC_FP64 0
#define MKG_PREC_FP32 1 /* FP32 populations stored as deviations p_i - w_i * ref */

/* ---- face boundary codes (faces ordered x-, x+, y-, y+, z-, z+) ---- */
#define MKG_BC_PERIODIC 0
#define MKG_BC_BOUNCE 1    /* flow: half-way bounce-back, moving wall / inlet with u_face;
                              scalar: zero flux */
#define MKG_BC_DIRICHLET 2 /* scalar only: anti-bounce-back to value s_face */
#define MKG_BC_OUTFLOW 3   /* unknown incoming populations copied from the cell itself */
#define MKG_BC_PRESSURE 4  /* flow only: anti-bounce-back to the equilibrium at density rho0 and
                              the cell's velocity (pressure outlet) */
#define MKG_BC_VELOCITY 5  /* flow only: velocity face by non-equilibrium extrapolation
                              (mkFluidsUser::boundary_vel, MaterKit/src/mkFluidsUser.cpp:
                              455-505; bdyVelocity, MaterKit/inc/mkFluidsLBM.h:63): after the
                              collision, each cell of the face layer takes the post-collision
                              populations of its inward neighbour n plus f_eq(rho_n, u_b) -
                              f_eq(rho_n, u_n), u_n = sum c f*_n / rho_n, u_b the face velocity;
                              populations entering the face layer from outside are its own
                              (zero gradient).  Two velocity faces must not share an edge. */

/* ---- anisotropy forms a_s(n), n the unit normal in the crystal frame ---- */
#define MKG_ANISO_CUBIC 0 /* 1 - 3 eps4 + 4 eps4 sum n^4 (Karma-Rappel), eps4 */
#define MKG_ANISO_CBC 1   /* 1 + e1 (Q - 3/5) + e2 (3 Q + 66 S - 17/7), Q = sum n^4,
                             S = nx^2 ny^2 nz^2 (two cubic harmonics: fcc / bcc / cubic);
                             aniso = {e1, e2} */
#define MKG_ANISO_HCP 2   /* 1 + e6 Re (nx + i ny)^6 + ez (3 nz^2 - 1) / 2 (six-fold basal
                             anisotropy and the c axis: hcp, snowflakes); aniso = {e6, ez} */
#define MKG_ANISO_FACET 3 /* 1 + ef (sum sqrt(n_k^2 + nu^2) - sqrt(1 + nu^2) - 2 nu): cusps at
                             the {100} facets rounded over nu; aniso = {ef, nu} */

/* ---- lattice masks for mkg_step ---- */
#define MKG_FLOW 1
#define MKG_PHASE 2
#define MKG_THERMAL 4
#define MKG_SOLUTE 8

/* ---- field ids for get/set ---- */
#define MKG_F_POP_FLOW 0    /* Q x N  pre-collision f_i at the current time level */
#define MKG_F_POP_PHASE 1   /* Qg x N */
#define MKG_F_POP_THERMAL 2 /* 7 x N */
#define MKG_F_POP_SOLUTE 3  /* 7 x N */
#define MKG_F_RHO 4         /* N */
#define MKG_F_VEL 5         /* 3 x N, Guo half-force velocity */
#define MKG_F_TEMP 6        /* N */
#define MKG_F_CONC 7        /* N, the conserved solute variable c = A(phi) U */
#define MKG_F_PHI 8         /* N, phase field at the last phase moment pass */
#define MKG_F_PHI_PREV 9    /* N */
#define MKG_F_FLAG 10       /* N int8 */
#define MKG_F_GRAIN 11      /* N int16 */
#define MKG_F_KAPPA 12      /* N, per-cell conductivity multiplier */

#define MKG_DT_F64 0
#define MKG_DT_F32 1
#define MKG_DT_I8 2
#define MKG_DT_I16 3

/* ---- reductions (FP64 on the device) ---- */
#define MKG_R_MASS 0      /* sum rho */
#define MKG_R_ENERGY 1    /* sum T */
#define MKG_R_SOLUTE 2    /* sum c */
#define MKG_R_SOLID 3     /* mean solid fraction (1+phi)/2 */
#define MKG_R_PHI 4       /* sum phi */
#define MKG_R_NUSSELT 5   /* 1 + <u_z T>/(alpha dT / H) for z-vertical Rayleigh-Benard */
#define MKG_R_MOMENTUM 6  /* out[0..2] = sum rho u */
#define MKG_R_UMAX 7      /* max |u| */
#define MKG_R_NONFINITE 8 /* number of non-finite populations */
#define MKG_R_TIP 9       /* out[0..2]: furthest phi = 0 crossing (interpolated between a cell
                             with phi >= 0 and its +x / +y / +z neighbour) from the domain
                             centre along +x, +y, +z: the dendrite tips; -1e300 if none */

#define MKG_MAX_GRAINS 64

typedef struct mkg_params {
    int32_t nx, ny, nz;   /* cells; linear index (x*ny + y)*nz + z, z contiguous */
    int32_t flow_q;       /* 0 (off), 19 or 27 */
    int32_t phase_q;      /* 0 (off), 7 or 19 */
    int32_t thermal;      /* 0/1: D3Q7 thermal ADE */
    int32_t solute;       /* 0/1: D3Q7 solutal ADE */
    int32_t precision;    /* MKG_PREC_* */
    int32_t flow_op;      /* MKG_OP_* (informational; the rates below define it) */
    int32_t pf_coupling;  /* phase-field driving variable: 0 none, 1 temperature, 2 solute U */
    int32_t ngrains;      /* 0: no rotation; else per-cell grain ids index rot[] (1-based) */
    int32_t has_kappa;    /* per-cell thermal conductivity multiplier present */
    int32_t bc_f[6], bc_g[6], bc_h[6], bc_c[6]; /* per lattice, per face */
    int32_t flow_incomp;  /* 1: He-Luo incompressible equilibrium (rho0 in the u terms,
                             u = (j + F/2) / rho0; MaterKit/inc/mkNode.h getEquPDFsHe) */
    /* flow */
    double s_nu;          /* shear rate = 1/tau */
    double s_bulk;        /* bulk (trace) rate */
    double s_odd;         /* odd ghost-moment rate (TRT: omega-) */
    double s_even;        /* even ghost-moment rate */
    double rho0;
    double force[3];      /* constant body force density */
    double buoy[3];       /* buoyancy: F += rho0 * buoy * (T - T_ref) */
    double T_ref;
    double u_face[6][3];  /* wall / inlet velocity per face */
    /* thermal */
    double alpha_l, alpha_s; /* lattice diffusivities in liquid / solid */
    double latent_T;         /* source w_i * latent_T * (phi^n - phi^{n-1}) */
    double T_face[6];
    /* solute */
    double D_l, D_s, k_part, latent_C, q_min;
    double at_coef;       /* anti-trapping coefficient a_t (Karma: 1/(2 sqrt 2)); 0 = off */
    double C_face[6];
    /* phase field (Karma-Rappel units W0 = tau0 = 1 scaled by dx, dt) */
    double pf_dt;         /* dt / tau0 */
    double pf_wdx2;       /* W0^2 / dx^2 */
    double eps4;
    double lambda;
    double phi_s;         /* flag threshold: phi >= phi_s solid, > -phi_s interface */
    int32_t aniso_kind;   /* MKG_ANISO_*: the form of a_s(n) (relaxation_anisotropy_cbc / _hcp,
                             anisotropy_facet; MaterKit/inc/mkPhaseField.h:49-64) */
    int32_t pf_model;     /* MKG_PF_KARMA_RAPPEL or MKG_PF_ALLEN_CAHN (see below) */
    double aniso[2];      /* parameters of the non-cubic forms (see MKG_ANISO_*) */
    /* deviation references for FP32 storage */
    double ref_f, ref_g, ref_h, ref_c;
    double rot[MKG_MAX_GRAINS][9]; /* row-major lab<-crystal rotation per grain */
    /* rotating walls (bdyRotation / bdySpiral, MaterKit/inc/mkFluidsLBM.h:64-65; angVel_0,
       :180): the velocity of a bounce or velocity face at the wall point r is
       u_face + the tangential part of wall_omega x (r - wall_center), lattice units, with
       r = the cell centre moved half a cell through the face */
    double wall_omega[3];
    double wall_center[3];
} mkg_params;

/* Phase-field model of the phase lattice.
 * MKG_PF_KARMA_RAPPEL  anisotropic dendrite growth (relaxation_anisotropy_thermal / _solute,
 *                      MaterKit/inc/mkPhaseField.h:41-47); cells are re-classified from phi
 *                      (reset_flags) and solid cells act on the flow.
 * MKG_PF_ALLEN_CAHN    conservative Allen-Cahn interface (relaxation_AllenCahn and
 *                      relaxation_AllenCahn_growth_thermal, MaterKit/inc/mkPhaseField.h:39-40,
 *                      run by run_flow_AllenCahn, MaterKit/MaterKit.cpp:186-242): the
 *                      counter-term (1 - omega_g/2) w_i (c_i . n)(1 - phi^2) dx/(sqrt 2 W0) of
 *                      Chai et al. (2018) with n = s/(|s| + 1e-12), s = sum_i c_i w_i phi(x+c_i),
 *                      plus the growth source -lambda u (1 - phi^2)^2 dt when pf_coupling
 *                      selects u; flags are not re-classified (the interface is carried by the
 *                      flow, not a solid). */
#define MKG_PF_KARMA_RAPPEL 0
#define MKG_PF_ALLEN_CAHN 1

typedef struct mkg_ctx* mkg_handle;

int mkg_abi_version(void);
size_t mkg_params_size(void);
const char* mkg_last_error(void);

/* Fill p with neutral defaults (everything off, periodic, rho0 = 1, phi_s = 0.99). */
int mkg_params_default(mkg_params* p);

/* Read a BoltzTechs XML case (the reference schema plus an optional <solver> block) and derive
 * lattice-unit parameters.  run_info (optional, 4 ints): numLoops, numFrames, seed count, 0. */
int mkg_params_from_xml(const char* path, mkg_params* p, int32_t* run_info);

/* Seeds of a case file: explicit <phasechange><seed> entries (ZYX Euler angles) or
 * <solver><seedsRandom> count radius rngseed (uniform quaternions).  Writes up to max
 * seeds -- centres (3n), radii (n), row-major rotations (9n, may be NULL) -- and returns the
 * count, or -1. */
int mkg_seeds_from_xml(const char* path, int max, double* centers, double* radii, double* rot);

int mkg_create(const mkg_params* p, mkg_handle* out);
int mkg_destroy(mkg_handle h);
int mkg_get_params(mkg_handle h, mkg_params* p);
/* Update scalar parameters (rates, forces, eps4, lambda...) of a live handle; grid, lattices
 * and precision must not change. */
int mkg_set_params(mkg_handle h, const mkg_params* p);

/* Initial state from macroscopic fields (host arrays of N doubles, NULL = uniform default):
 * equilibrium populations, phase buffers, flags from phi (walls kept). */
int mkg_init_equilibrium(mkg_handle h, const double* rho, const double* vel3, const double* T,
                         const double* C, const double* phi);

/* Seeds: centres (3*n, lattice units), radii (n), grain per seed = index+1.  Sets phi
 * (tanh profile, interface width W0/dx), grain ids (nearest seed surface), flags. */
int mkg_init_seeds(mkg_handle h, int n, const double* centers, const double* radii);

/* Random seed placement for multi-dendrite cases: uniform centres in [r, dim-1-r] and uniform
 * unit quaternions (Shoemake), both drawn from a splitmix64 stream of `seed`. */
int mkg_random_seeds(uint64_t seed, int n, double radius, int nx, int ny, int nz,
                     double* centers, double* quats);
/* Anisotropy of one crystal at an interface normal, the query of mkPhaseField::anisotropy /
 * anisotropy_facet / anisotropy_cbc / anisotropy_hcp (MaterKit/inc/mkPhaseField.h:61-64) and
 * of the per-grain mkMultiphaseField versions (MaterKit/inc/mkMultiphaseField.h:31-34), with
 * the forms of the kernels: kind MKG_ANISO_*, eps4 and par[2] its parameters, rot the
 * row-major lab <- crystal rotation (NULL: identity), normal the lab-frame normal (any length).
 * Returns a_s(n) and n_v = R a_s (D - n (n.D)), D = d a_s / d n in the crystal frame, so that
 * the flux of the phase equation is J = |grad phi| (a_s^2 n + n_v).  Returns 1 (a_s = 1,
 * n_v = 0) for a zero normal, 0 otherwise. */
int mkg_aniso_eval(int kind, double eps4, const double* par, const double* rot,
                   const double* normal, double* a_s, double* n_v);

/* Streaming rate of the device in GB/s (read + write) of a plain copy kernel over two
 * buffers of `bytes` each (allocated and freed by the call), averaged over `reps` launches:
 * the practical ceiling against which the bandwidth-bound kernels are read.  No reference
 * counterpart (a measurement aid). */
int mkg_stream_copy_rate(size_t bytes, int reps, double* gbs);

/* Quaternion (w,x,y,z) to row-major rotation matrix. */
int mkg_quat_to_rot(const double* q, double* R);

/* Run nsteps time steps advancing the lattices in mask (others are read for coupling only).
 * A single fused kernel pair per step is used when all advanced lattices share a parity. */
int mkg_step(mkg_handle h, int nsteps, int mask);
int mkg_sync(mkg_handle h);
int64_t mkg_step_count(mkg_handle h, int lattice_bit);

int mkg_get_field(mkg_handle h, int field, void* host, int dtype, size_t count);
int mkg_set_field(mkg_handle h, int field, const void* host, int dtype, size_t count);

int mkg_reduce(mkg_handle h, int which, double* out);

/* The same reductions (same values, bit for bit) without any device allocation: no
 * counterpart in the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The kernels read
 * the cell values straight from the population storage and keep their per-block partials in
 * one fixed device-resident array per GPU (4 MiB, held until the process ends).  Grids of
 * more than 2^25 cells (33,554,432) are rejected with an error.  Calls on one GPU share that
 * array: calls that overlap on one GPU (other threads, other handles' streams) race. */
int mkg_reduce_noalloc(mkg_handle h, int which, double* out);

int mkg_checkpoint_save(mkg_handle h, const char* path);
int mkg_checkpoint_load(mkg_handle h, const char* path);

/* Device time (ms) of the last mkg_step call, per kernel class: out[0] = phase moment pass,
 * out[1] = fused collide-stream, out[2] = total; out[3] = number of steps timed. */
int mkg_last_timing(mkg_handle h, double* out);
int mkg_enable_timing(mkg_handle h, int on);

/* Discrete adjoint (no counterpart in the reference) of nsteps fused phase-field (D3Q19) +
 * thermal (D3Q7) steps from the current state, with or without a D3Q19 flow (the coupled step
 * of run_flow_thermal_dendrite, MaterKit/MaterKit.cpp:500-518: the melt advects both scalars,
 * feels the Boussinesq buoyancy and sees the solid through partial bounce-back; compressible
 * equilibrium, periodic / bounce / outflow flow faces): a forward pass with checkpoints every
 * ckpt steps (the last segment writes its tape as it steps), the FP64 loss
 * L = 1/2 sum (phi_N - phi_target)^2 with phi_N = sum_i g_i after the last step, then the
 * reverse sweep with per-segment recompute of the earlier segments (segment s - 1 recomputed
 * while segment s is reversed on a second stream).  ckpt = 0 keeps a tape instead: the
 * forward pass stores what the reverse sweep reads for every step (phi^n and the thermal
 * populations, 8 values per cell and step, plus the flow populations in FP64) and nothing is
 * recomputed.  Outputs L, dL/deps4 and (optional, needs has_kappa) dL/dkappa per cell (N
 * doubles); times (optional, 4 doubles): forward ms, recompute ms not overlapped with the
 * reverse, reverse phase ms, total.  The handle is left at the final forward state. */
int mkg_adjoint_run(mkg_handle h, int nsteps, int ckpt, const double* phi_target, double* loss,
                    double* dL_deps4, double* dL_dkappa, double* times);

/* Discrete adjoint of nsteps flow-only steps (no phase, thermal or solute lattice) from the
 * current state: the generalised moment collision (BGK, TRT, MRT) with Guo forcing and
 * streaming with periodic, bounce (any wall velocity) or outflow faces and wall cells
 * (the adjoints of mkFluidsLBM::propagate / relaxation and the MRT overrides,
 * MaterKit/inc/mkFluidsLBM.h:47-51, mkFluidsMRT.h:42-46; no counterpart in the reference).
 * Loss L = 1/2 sum over non-wall cells |u_N - u_target|^2 (u_target: 3 x N, cell-interleaved;
 * u the Guo half-force velocity), reduced in FP64.  grad[7] = dL/d(force x, y, z) and
 * dL/d(s_nu, s_bulk, s_odd, s_even), each rate taken as independent (a BGK gradient with
 * respect to its single rate is the sum of the four).  The handle is left at step nsteps.
 * The forward step writes every step's populations to a tape (Q N values of the storage type
 * per step). */
int mkg_flow_adjoint_run(mkg_handle h, int nsteps, const double* u_target, double* loss,
                         double* grad);

/* ---- Multi-phase field ---------------------------------------------------------------
 * n order parameters phi_j in [-1, 1], each with its own D3Q19 population lattice, plus an
 * optional D3Q7 thermal lattice: the per-grain phase lattices (gVec, phiVec) of
 * mkFluidsUser::propagate / relaxation (MaterKit/src/mkFluidsUser.cpp:68-134, 136-299; the
 * node state at MaterKit/inc/mkNode.h:158-164, phase_number at MaterKit/inc/mkMultiphaseField.h:56)
 * as driven by run_multiPhaseField (MaterKit/MaterKit.cpp:529-576).  Each order parameter
 * follows a conservative Allen-Cahn LB (Chai et al. 2018, the getAddedPhasePDFs term) with
 * its own advection velocity and the pairwise interaction term of the template; the combined
 * field phi = (n - 1) + sum_j phi_j is kept beside them.  One handle owns its device memory
 * and one HIP stream.
 *
 * MKG_MPF_KARMA_RAPPEL makes every order parameter an anisotropic Karma-Rappel grain of its
 * own form (MKG_ANISO_*), strength and orientation, driven by T or by the solute's U, with the
 * pairwise interaction on the same time scale tau0 a_s^2 -- the per-grain anisotropy and the
 * thermal / solutal growth of mkMultiphaseField (anisotropy, anisotropy_facet, anisotropy_cbc,
 * anisotropy_hcp, relaxation_anisotropy_thermal, relaxation_solute, relaxation_thermal;
 * MaterKit/inc/mkMultiphaseField.h:31-45).  One grain is the single-grain phase field of
 * mkg_step exactly.  The optional D3Q7 solute lattice follows the single-grain solute equation
 * (c = A(phi) U, anti-trapping current) with the combined phi. */
#define MKG_MPF_MAX 4
typedef struct mkg_mpf_params {
    int32_t nx, ny, nz;
    int32_t n;                  /* order parameters, 1..MKG_MPF_MAX */
    int32_t precision;          /* MKG_PREC_FP64 or MKG_PREC_FP32 (deviation storage) */
    int32_t thermal;            /* 1: D3Q7 thermal lattice */
    int32_t bc[6], bc_h[6];     /* face codes of the phase lattices / the thermal lattice */
    double dt;                  /* dt / tau0 */
    double wdx;                 /* W0 / dx */
    double u[MKG_MPF_MAX][3];   /* advection velocity of each order parameter (template:
                                   0.02 in x for j = 1, MaterKit/src/mkFluidsUser.cpp:219-230) */
    double kappa[MKG_MPF_MAX];  /* interaction strength (template: 0 for j = 0, else 1, :250-255) */
    double growth;              /* coefficient of the growth term (template: 0, :266-271) */
    double lambda;              /* coupling constant of the growth term */
    double eps;                 /* normal regularisation n = s / (|s| + eps) */
    double alpha, latent_T, T_ref;
    double T_face[6];
    double phi_face[6];         /* Dirichlet value of the phase lattices */
    int32_t model;              /* MKG_MPF_TEMPLATE or MKG_MPF_KARMA_RAPPEL */
    int32_t coupling;           /* field driving Karma-Rappel grains: 0 none, 1 T, 2 U */
    int32_t aniso_kind[MKG_MPF_MAX];  /* MKG_ANISO_* form of each grain */
    double eps4[MKG_MPF_MAX];         /* cubic strength of each grain (MKG_ANISO_CUBIC) */
    double aniso[MKG_MPF_MAX][2];     /* parameters of the other forms, per grain */
    double rot[MKG_MPF_MAX][9];       /* row-major lab <- crystal rotation of each grain */
    int32_t solute;             /* 1: D3Q7 solute lattice (Karma-Rappel grains) */
    int32_t bc_c[6];            /* face codes of the solute lattice */
    double D_l, D_s, k_part, q_min, latent_C, at_coef;  /* as in mkg_params */
    double C_face[6];           /* Dirichlet U of the solute faces */
} mkg_mpf_params;

#define MKG_MPF_TEMPLATE 0      /* conservative Allen-Cahn order parameters (the template) */
#define MKG_MPF_KARMA_RAPPEL 1  /* anisotropic grains, see above */

typedef struct mkg_mpf_ctx* mkg_mpf_handle;

#define MKG_MPF_POPS 0      /* n x 19 x N pre-collision populations (canonical, doubles) */
#define MKG_MPF_PHI 1       /* (n + 1) x N: phi_j at the last moment pass, then combined phi */
#define MKG_MPF_PHI_PREV 2  /* (n + 1) x N: the pass before */
#define MKG_MPF_THERMAL 3   /* 7 x N thermal populations (doubles) */
#define MKG_MPF_FLAG 4      /* N int8: 0 fluid, 3 wall (half-way bounce-back) */
#define MKG_MPF_SOLUTE 5    /* 7 x N solute populations (doubles) */

size_t mkg_mpf_params_size(void);
int mkg_mpf_params_default(mkg_mpf_params* m);
int mkg_mpf_create(const mkg_mpf_params* m, mkg_mpf_handle* out);
int mkg_mpf_destroy(mkg_mpf_handle h);
/* Equilibrium start: phi n x N order parameters, T N (NULL: T_ref), both doubles; the
 * solute lattice (if any) starts at c = 0. */
int mkg_mpf_init(mkg_mpf_handle h, const double* phi, const double* T);
/* The same with the solute c (N doubles, NULL: 0). */
int mkg_mpf_init_fields(mkg_mpf_handle h, const double* phi, const double* T, const double* C);
/* nsteps steps: phase moment pass, then collide-stream of every lattice (fused). */
int mkg_mpf_step(mkg_mpf_handle h, int nsteps);
int mkg_mpf_sync(mkg_mpf_handle h);
/* Read / write a field (MKG_MPF_*); count is the element count of the host buffer. */
int mkg_mpf_get(mkg_mpf_handle h, int which, void* host, size_t count);
int mkg_mpf_set(mkg_mpf_handle h, int which, const void* host, size_t count);
/* FP64 sums over the non-wall cells of phi_j (n values) and of the combined phi. */
int mkg_mpf_sums(mkg_mpf_handle h, double* out);

/* ---- Finite-volume phase field and solute ----------------------------------------------
 * Explicit cell-centred forms of mkPhaseField::evolution_anisotropy
 * (MaterKit/inc/mkPhaseField.h:51, "on FVM") and of mkSolutalField::evolution_ATC1 /
 * evolution_ATC2 (MaterKit/inc/mkSolutalField.h:44-46), the solute update of run_solutal_only
 * (MaterKit/MaterKit.cpp:689-741).  Fields: phi and c = A(phi) U (A = ((1+k) - (1-k) phi)/2),
 * N values each.  A step: phi' = phi + dt/a^2 (W^2/dx^2 div J + Q) with the face fluxes and
 * anisotropy of the lattice solver, then c' = c - sum_faces J_out + (phi' - phi)/2 with the
 * diffusive flux D q(phi) grad U and the anti-trapping current.  The handle takes the phase and
 * solute parameters of an mkg_params block (pf_dt, pf_wdx2, eps4, lambda, aniso_*, D_l, D_s,
 * k_part, at_coef, q_min; faces bc_g: periodic or bounce = zero flux; bc_c: periodic, bounce or
 * Dirichlet with C_face as the face value of U); ngrains must be 0.
 * With thermal = 1 the potential is the temperature (mkSolutalField::evolution_thermal,
 * MaterKit/inc/mkSolutalField.h:47-48, "potential is temperature"): the phase source takes
 * u = T and T' = T - sum_faces J_out + L_T (phi' - phi) with J_out = -alpha_f (T_nb - T),
 * alpha_f the mean of alpha(phi) = alpha_l (1 - f_s) + alpha_s f_s of the two cells (faces
 * bc_h, Dirichlet values T_face; 0 < alpha <= 1/6); fields MKG_FVM_PHI and MKG_FVM_TEMP, and
 * mkg_fvm_sums gives the sums of phi and T. */
typedef struct mkg_fvm_ctx* mkg_fvm_handle;
#define MKG_FVM_PHI 0   /* N, phase field */
#define MKG_FVM_CONC 1  /* N, c = A(phi) U */
#define MKG_FVM_U 2     /* N, U = c / A(phi) (get only) */
#define MKG_FVM_TEMP 3  /* N, T: the potential of a thermal handle */

int mkg_fvm_create(const mkg_params* p, mkg_fvm_handle* out);
int mkg_fvm_destroy(mkg_fvm_handle h);
/* Set a field (MKG_FVM_PHI or MKG_FVM_CONC, N doubles) at the current level. */
int mkg_fvm_set(mkg_fvm_handle h, int which, const double* host, size_t count);
/* nsteps explicit steps; evolve_phase 0 holds phi fixed (run_solutal_only). */
int mkg_fvm_step(mkg_fvm_handle h, int nsteps, int evolve_phase);
int mkg_fvm_sync(mkg_fvm_handle h);
int mkg_fvm_get(mkg_fvm_handle h, int which, double* host, size_t count);
/* FP64 sums of phi and of c (or T) over every cell (out[0], out[1]). */
int mkg_fvm_sums(mkg_fvm_handle h, double* out);

/* ---- Two-phase flow ----------------------------------------------------------------------
 * Phase-field two-phase flow of mkFluidsMultiphase (MaterKit/inc/mkFluidsMultiphase.h:20-99;
 * run_flow_multiphase, MaterKit/MaterKit.cpp:290-331): a Cahn-Hilliard lattice for the order
 * parameter phi in [0, 1] and a velocity-pressure flow lattice, both D3Q19, in the model of
 * Liang et al. (2014) whose parameters the reference class carries (rho_l, rho_g, sigma, W_0,
 * beta = 12 sigma/W, kappa = 1.5 sigma W, eta; mkFluidsMultiphase.h:81-97), with BGK
 * collisions: mu = 4 beta phi (phi-1)(phi-1/2) - kappa lap phi, surface force mu grad phi,
 * rho = rho_g + phi (rho_l - rho_g), nu = nu_g + phi (nu_l - nu_g), mobility
 * M = eta cs^2 (tau_h - 1/2).  Faces: periodic or bounce (no-slip wall, zero flux).
 * Wetting (initialize_wetting, mkFluidsMultiphase.h:35; theta, :95): a bounce face whose bit
 * is set in wet_faces carries the contact angle theta (degrees, measured in the phi = 1
 * fluid) through the cubic surface energy -sigma cos(theta) phi^2 (3 - 2 phi), i.e.
 * n.grad phi = (4/W) cos(theta) phi (1 - phi) along the outward normal n; the phi stencils
 * take the ghost value phi_e + (4/W) cos(theta) phi_e (1 - phi_e) across such a face (phi_e
 * the edge cell's value).  wet_faces = 0 leaves every wall neutral (90 degrees). */
typedef struct mkg_mp_params {
    int32_t nx, ny, nz, precision;
    int32_t bc[6];
    double rho_l, rho_g;      /* densities of phi = 1 and phi = 0 */
    double nu_l, nu_g;        /* kinematic viscosities (lattice units) */
    double sigma, width;      /* surface tension, interface width W (cells) */
    double mobility, tau_h;   /* M and the Cahn-Hilliard relaxation time (> 1/2) */
    double gravity[3];        /* body acceleration: force rho g */
    double force[3];          /* uniform body-force density */
    int32_t wet_faces;        /* bit f (order of bc[]): face f is a wetting wall */
    int32_t visc_model;       /* MKG_MP_VISC_*: viscosity across the interface */
    double contact_angle;     /* degrees, in (0, 180), on the wetting faces */
} mkg_mp_params;

/* Viscosity across the interface, at phi clamped to [0, 1]: nu = nu_g + phi (nu_l - nu_g);
 * mu = rho nu linear in phi; or 1/mu = phi/mu_l + (1 - phi)/mu_g (continuous shear stress
 * across a layered interface, the closest to the sharp-interface profile at large ratios). */
#define MKG_MP_VISC_NU_LINEAR 0
#define MKG_MP_VISC_MU_LINEAR 1
#define MKG_MP_VISC_MU_HARMONIC 2

typedef struct mkg_mp_ctx* mkg_mp_handle;
#define MKG_MP_POP_FLOW 0   /* 19 x N flow populations (canonical, doubles) */
#define MKG_MP_POP_PHASE 1  /* 19 x N Cahn-Hilliard populations */
#define MKG_MP_PHI 2        /* N order parameter (get) */
#define MKG_MP_MU 3         /* N chemical potential (get) */
#define MKG_MP_PRESSURE 4   /* N pressure p (get) */
#define MKG_MP_VEL 5        /* 3 x N velocity, cell-interleaved (get) */
#define MKG_MP_PHIU 6       /* 3 x N phi u of the last step, cell-interleaved (set) */

size_t mkg_mp_params_size(void);
int mkg_mp_params_default(mkg_mp_params* m);
int mkg_mp_create(const mkg_mp_params* m, mkg_mp_handle* out);
int mkg_mp_destroy(mkg_mp_handle h);
/* Equilibrium start from phi (N), u (3 x N, cell-interleaved, or NULL) and p (N or NULL). */
int mkg_mp_init(mkg_mp_handle h, const double* phi, const double* u, const double* p);
int mkg_mp_step(mkg_mp_handle h, int nsteps);
int mkg_mp_sync(mkg_mp_handle h);
int mkg_mp_get(mkg_mp_handle h, int which, double* host, size_t count);
int mkg_mp_set(mkg_mp_handle h, int which, const double* host, size_t count);
/* FP64 sum of phi over every cell. */
int mkg_mp_sums(mkg_mp_handle h, double* out);
/* Two-phase parameters from the <multiphases> block of a case file (run_flow_multiphase's
 * case, MaterKit/BoltzTechs - multiPhaseFlow.xml): densities relative to the rare fluid,
 * kinematic viscosities and mobility scaled by mk_dt/mk_dx^2, W0, sigma, force; the body
 * acceleration from flow/acceleration; faces from solver/bcFlow; multiphases/contactAngle
 * (default 90) on every bounce face (wet_faces) unless it is 90; solver/viscosityModel
 * (nu-linear, mu-linear, mu-harmonic; default nu-linear).  run_info (4, may be NULL):
 * numLoops, numFrames, seed count, start geometry (solver/twoPhaseInit: 0 "layered", the
 * dense fluid in the lower half along y; 1 "drops", one dense drop per seed). */
int mkg_mp_params_from_xml(const char* path, mkg_mp_params* m, int32_t* run_info);

#ifdef __cplusplus
}
#endif
#endif /* MKG_H */
