// Kernel argument blocks and launcher declarations shared by the C-ABI layer (api.cpp) and
// the HIP kernels (step.hip, aux.hip).  Plain structs passed by value to the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace mkg {

// One population lattice as seen by a kernel.
struct LatArgs {
    void* pop;        // slot 0 of the lattice; cell c of slot i at pop + i * sl_stride
                      // + pcell(c) (StepArgs: population addressing)
    void* face;       // face buffer, Q x F elements: slot i of face cell k at face + i*F + k
                      // (see face_slot)
    int32_t bc[6];    // face codes x-, x+, y-, y+, z-, z+
    int32_t npmask;   // bit f set when face f is not periodic
    int32_t lay;      // current AA layout: 0 = own slots hold f_i, 1 = shifted
    int32_t adv;      // advance this lattice in this launch
    uint32_t pg_stride;  // elements from one page of this lattice to the next (FP64 kernels;
                         // the FP32 kernels take StepArgs::pg_stride, every lattice sharing pages)
    int32_t fid;      // this lattice's entry of the boundary table (FaceTab): 0 flow, 1 phase
                      // (every order parameter), 2 thermal, 3 solute
    double ref;       // deviation reference: stored value = f_i - w_i * ref
    double fval[6];   // Dirichlet face values (scalar lattices)
};

// Boundary parameters of the face rules (face_in) in device memory, a copy of the same fields
// of a launch's StepArgs and lattices.  The face rules read them through StepArgs::ftab inside
// the face branches.  Read as kernel arguments they are loaded at kernel entry and held in
// scalar registers through the whole step, which spills scalar registers in the kernels with
// face rules.
struct FaceTab {
    int32_t bc[4][6];     // by LatArgs::fid
    int32_t wall_rot;
    int32_t unused;
    double fval[4][6];
    double ref[4];
    double u_face[6][3];
    double rho0;
    double wall_omega[3], wall_center[3];
};

struct StepArgs {
    int32_t nx, ny, nz;
    uint32_t N;       // number of cells
    uint32_t S;       // stride of the per-cell arrays (phi, flags, kappa): N rounded up
    // Population addressing: cell c of slot i of a lattice is at
    //   L.pop + i * sl_stride + (c >> pg_shift) * stride + (c & pg_mask),
    // stride = pg_stride in FP32 kernels, L.pg_stride in FP64 kernels.
    // Paged layout: pages of 2^pg_shift cells (sl_stride = page cells).  FP32: a page holds
    // every slot of every lattice of the handle, slot by slot (pg_stride = all slots x page
    // cells).  FP64: each lattice has pages of its own slots (L.pg_stride = its slots x page
    // cells), so that every lattice stays within 4 GiB of 32-bit byte offsets.
    // Slot-major layout: one array per slot (pg_shift = 31, pg_mask = 2^31 - 1,
    // sl_stride = S, strides 0).
    uint32_t pg_shift, pg_mask, pg_stride, sl_stride;
    uint32_t nblk;    // blocks in the launch (XCD-aware remap when divisible by 8)
    uint32_t F;       // face-buffer stride: cells on the six faces, 2 (ny nz + nx nz + nx ny)
    int32_t has_walls, has_kappa, ngrains, pf_coupling;
    int32_t zstage;   // O steps move the z populations through LDS (needs 256 % nz == 0)
    int32_t incomp;   // He-Luo incompressible equilibrium: rho0 in the u terms
    int32_t xcd_chunk;  // block order (remap_block): runs of this many blocks per XCD, 0 off
    LatArgs f, g, h, c;
    const void* phi_cur;   // storage type, phase field at time level n
    const void* phi_prev;  // level n-1
    int8_t* flag;
    const int16_t* grain;
    const void* kappa;     // storage type
    const double* rot;     // ngrains x 9, lab <- crystal
    double s_nu, s_bulk, s_odd, s_even, rho0;
    double force[3], buoy[3], T_ref;
    double u_face[6][3];
    double alpha_l, alpha_s, latent_T;
    double D_l, D_s, k_part, latent_C, q_min;
    double at_w;  // anti-trapping a_t * W0/dx
    double pf_dt, pf_wdx2, eps4, lambda, phi_s, omega_g;
    int32_t aniso_kind, pf_model;
    double aniso[2];
    double ac_zc;  // Allen-Cahn counter-term scale (1 - omega_g/2) / (sqrt 2 W0/dx)
    // rotating walls: wall velocity u_face + tangential part of wall_omega x (r - wall_center)
    int32_t wall_rot;  // wall_omega != 0
    int32_t nee_mask;  // bit f: face f of the flow lattice is a velocity face
    double wall_omega[3], wall_center[3];
    // velocity faces: the replaced post-collision populations of each face cell (storage type,
    // Q x F, slot i of face cell k at i * F + k), written by the pre-pass of the step
    void* nee;
    // adjoint tape of the step (phase field + thermal model): phi^n at tape[cell] and the
    // pre-collision thermal populations at tape[(1 + k) * S + cell]; null: no tape
    void* tape;
    // device copy of the boundary parameters above (kernels with face rules)
    const FaceTab* ftab;
};

// The boundary table of a launch's arguments (lattices f, g, h, c with fid 0..3).
FaceTab facetab_of(const StepArgs& a);
// A handle's device copy of its boundary table: allocated with the handle, uploaded where the
// handle's parameters change (upload waits for the copy on stream s).
hipError_t facetab_alloc(FaceTab** dev);
hipError_t facetab_upload(FaceTab* dev, const StepArgs& a, hipStream_t s);

// Population block of a handle: the slots of every lattice, then every lattice's Q x F face
// buffer, in one allocation of `total` elements.  Paged (pages of 2^pg_shift cells) when the
// kernels can address every lattice with 32-bit byte offsets from its slot 0: FP32 pages hold
// every slot of every lattice (the block below 4 GiB), FP64 pages the slots of one lattice
// (each lattice below 4 GiB); slot-major otherwise, or when MKG_PAGES=0.  MKG_PAGE_SHIFT sets
// the page size (default 2^13 cells).
constexpr int kMaxLattices = 8;
struct PopBlock {
    int paged = 0;
    uint32_t pg_shift = 31, pg_mask = 0x7FFFFFFFu, pg_stride = 0, sl_stride = 0;
    size_t total = 0;                   // elements
    size_t slot0[kMaxLattices] = {};    // element offset of slot 0 of lattice k
    size_t face0[kMaxLattices] = {};    // element offset of the face buffer of lattice k
    uint32_t lat_stride[kMaxLattices] = {};  // page stride of lattice k (LatArgs::pg_stride)
    void apply(StepArgs& a) const {
        a.pg_shift = pg_shift;
        a.pg_mask = pg_mask;
        a.pg_stride = pg_stride;
        a.sl_stride = sl_stride;
    }
    // lattice k of a block allocated at base (element size esz)
    void lattice(LatArgs& L, void* base, int k, size_t esz) const {
        L.pop = (char*)base + slot0[k] * esz;
        L.face = (char*)base + face0[k] * esz;
        L.pg_stride = lat_stride[k];
    }
};
// q[k]: slots of lattice k (0: absent); N cells, S per-cell array stride, F face cells.
PopBlock plan_pop_block(const int* q, int nlat, uint32_t N, uint32_t S, uint32_t F, size_t esz);

// Block order of the launches (StepArgs::xcd_chunk) for elements of esz bytes:
// MKG_XCD_CHUNK, else the default of that element size.
int xcd_chunk_setting(size_t esz);

// Selector of a fused-step instantiation: presence of each lattice.
struct Combo {
    int fq, pq, th, so;
};

// Returns 0 on success, -1 when the combination has no instantiation.
int launch_step(const Combo& k, int fp32, int odd, const StepArgs& a, hipStream_t s);
// Pre-pass of a step with velocity faces: the post-collision populations every face cell of a
// velocity face takes (non-equilibrium extrapolation from its inward neighbour), into a.nee.
int launch_nee(const Combo& k, int fp32, const StepArgs& a, hipStream_t s);
int launch_phase_moments(int pq, int fp32, int odd, const StepArgs& a, void* phi_out,
                         hipStream_t s);

// Canonical (layout-independent) pre-collision populations of one lattice into a dense
// Q x N buffer of doubles (full values, deviation reference added back).
// kind: 0 flow, 1 scalar.
int launch_canonical(int q, int fp32, const StepArgs& a, const LatArgs& L, int kind,
                     double* out, hipStream_t s);
// Write full-value populations (Q x N doubles) into own-slot layout.
int launch_store_pops(int q, int fp32, const StepArgs& a, const LatArgs& L, const double* in,
                      hipStream_t s);

// Macroscopic fields read straight from the AA storage (each lattice at its own layout):
// rho, u (3xN, Guo half-force velocity), T, C; any output may be null.
int launch_macro(int fq, int fp32, const StepArgs& a, double* rho, double* u, double* T,
                 double* C, hipStream_t s);

// Storage-type <-> double conversions of N-element scalar arrays.
int launch_to_double(int fp32, const void* src, double* dst, uint32_t n, hipStream_t s);
int launch_from_double(int fp32, const double* src, void* dst, uint32_t n, hipStream_t s);

// Equilibrium populations from macroscopic double arrays (own-slot layout); phase buffers
// phi0 / phi1 and flags from phi when given.
int launch_init_eq(int fp32, int fq, int pq, const StepArgs& a, const double* rho,
                   const double* u, const double* T, const double* C, const double* phi,
                   void* phi0, void* phi1, hipStream_t s);

// Seed geometry: phi (both buffers and the double scratch phid), grain ids, flags, phase
// populations.
int launch_init_seeds(int fp32, int pq, const StepArgs& a, int n, const double* centers,
                      const double* radii, double width, double ts, void* phi0, void* phi1,
                      double* phid, int16_t* grain, hipStream_t s);

// FP64 block reductions over double fields (null when absent); deterministic two-pass
// tree (per-block partials, then one block in fixed order).  out receives nout doubles.
int launch_reduce(int which, const StepArgs& a, const double* rho, const double* u,
                  const double* T, const double* C, const double* phi, double* partials,
                  double* out, int* nout, hipStream_t s);
// Count of non-finite values in a storage-type buffer of n elements.
int launch_nonfinite(int fp32, const void* p, size_t n, double* partials, double* out,
                     hipStream_t s);

// Arguments of the reverse (adjoint) step of the phase-field + thermal model.
struct AdjArgs {
    const void* gin;     // Q x N adjoint of the phase populations at level n+1 (partial)
    const void* nin;     // N: cell-wide phase adjoint still to be added to every population
    const void* hin;     // 7 x N adjoint of the thermal populations at level n+1
    const void* pin;     // N: adjoint of prev at level n+1
    void* gout;
    void* hout;
    void* pout;
    void* P;             // stencil contributions: 19 x N per direction, or 9 x N column sums
    void* kap_grad;      // N accumulated dL/dkappa (storage type)
    const void* phin;    // N: phi^n of the forward step being reversed
    const void* htape;   // pre-collision thermal populations of level n: slot k at htape + k*S
    double* eps_part;    // per-block FP64 partials of dL/deps4 (accumulated)
    // coupled model (flow + thermal + phase):
    const double* ftape; // Q x N canonical flow populations at level n (full values, FP64)
    const void* fin;     // Q x N adjoint of the flow populations at level n+1
    void* fout;          // Q x N adjoint of the flow populations at level n
    void* ut;            // 3 x N dL/d(velocity the scalars are advected with), collide -> flow
    const void* hcw_in;  // N cell-wide thermal adjoint of level n+1 (the buoyancy term)
    void* hcw_out;       // N cell-wide thermal adjoint of level n
    void* nacc;          // N cell-wide phase adjoint of level n: the flow kernel adds its part
};

// grouped: the launch's blocks hold whole z rows; the stencil contributions leave as nine
// column sums per cell for launch_adj_qsum (else 19 per-direction values for
// launch_adj_stencil).
int launch_adj_collide(int fp32, int grouped, const StepArgs& a, const AdjArgs& j,
                       hipStream_t s);
// Coupled model: the scalar/phase reverse with the advection velocity (flow 1), and the flow
// reverse with partial bounce-back, buoyancy and the scalars' velocity adjoint.
int launch_adj_collide_flow(int fp32, int grouped, const StepArgs& a, const AdjArgs& j,
                            hipStream_t s);
int launch_cadj_flow(int fp32, const StepArgs& a, const AdjArgs& j, hipStream_t s);
int launch_adj_stencil(int fp32, const StepArgs& a, const void* P, void* nout, hipStream_t s);
int launch_adj_qsum(int fp32, const StepArgs& a, const void* Q, void* nout, hipStream_t s);
int launch_adj_loss(int fp32, const StepArgs& a, const double* target, void* nseed, double* part,
                    double* out, hipStream_t s);
int launch_sum_partials(const double* part, uint32_t n, double* out, hipStream_t s);

// Flow adjoint (fadj.hip): loss and final adjoint from the canonical FP64 populations fN
// (partials 0..3 of an 8 x nblk array), and one reverse step over the tape of its level (the
// pre-collision populations the step kernel wrote: storage type, deviations, slot i at
// tape + i*S; partials 1..7).  The adjoint buffers hold the storage type (float when fp32).
int launch_fadj_loss(int q, int fp32, const StepArgs& a, const double* fN, const double* ut,
                     void* lam, double* part, hipStream_t s);
int launch_fadj_step(int q, int fp32, const StepArgs& a, const void* tape, const void* lam_in,
                     void* lam_out, double* part, hipStream_t s);

// Reductions without allocation (reduce_static.hip): per-block partials and the result live
// in a fixed device-resident array.  noalloc_prepare looks its address up on device dev (once
// per device); reduce_noalloc runs reduction `which` and copies out[0..3] to host.  macro:
// 0 the phase field only, 1 the scalar lattices, 2 flow and scalars.
int noalloc_prepare(int dev, std::string* err);
int reduce_noalloc(int dev, int fq, int fp32, const StepArgs& a, int which, int macro,
                   void* const* pops, const size_t* sizes, int npops, hipStream_t s,
                   double* host, std::string* err);

}  // namespace mkg
