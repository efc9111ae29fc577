// Explicit finite-volume solvers of the phase field (mkPhaseField::evolution_anisotropy,
// MaterKit/inc/mkPhaseField.h:51, "on FVM") and of the dilute-alloy solute with the
// anti-trapping current (mkSolutalField::evolution_ATC1 / evolution_ATC2,
// MaterKit/inc/mkSolutalField.h:44-46; run_solutal_only, MaterKit/MaterKit.cpp:689-741).
//
// One step is two launches over cell-centred fields (no populations):
//   k_fvm_phase   phi' = phi + dt/a^2 (W^2/dx^2 div J + Q): the face fluxes, anisotropy and
//                 source of the lattice solver's phase_source (device.h) plus the 7-point
//                 Laplacian that the lattice's diffusion would otherwise carry
//   k_fvm_solute  c' = c - sum_faces J_out + (phi' - phi)/2 with c = A(phi) U,
//                 J_out = -D q_f (U_nb - U) + s j_at,a; j_at = -a_t W/dx (1 + (1-k) U_f) dphi_f
//                 g/|g| on the face (g: normal difference, transverse means of phi)
// Both read the old fields and write the other buffer of a pair; c - phi/2 is conserved.
// With the temperature as the potential (mkSolutalField::evolution_thermal,
// MaterKit/inc/mkSolutalField.h:47-48) the phase source takes u = T and the second launch is
//   k_fvm_thermal T' = T - sum_faces J_out + L_T (phi' - phi), J_out = -alpha_f (T_nb - T),
//                 alpha_f the mean of the two cells' alpha(phi); T - L_T phi is conserved.
// When a block can hold whole z rows, the marching forms (k_fvm_phase_march,
// k_fvm_solute_march) evaluate each face once for both of its cells where they can.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/mkg.h"
#include "device.h"

namespace mkg {

void set_error(const std::string& msg);

template <class R>
__device__ inline R fvm_q(const StepArgs& a, R phi) {
    R q = (R)0.5 * ((R)1 - phi);
    if (a.D_l > 0) q += (R)(a.D_s / a.D_l) * (R)0.5 * ((R)1 + phi);
    return q < (R)a.q_min ? (R)a.q_min : q;
}
template <class R>
__device__ inline R fvm_A(const StepArgs& a, R phi) {
    return (R)0.5 * (((R)1 + (R)a.k_part) - ((R)1 - (R)a.k_part) * phi);
}
// The potential u that drives the phase source: U = c / A(phi) of the solute, or T itself
// (pf_coupling 1).
template <class R>
__device__ inline R fvm_u(const StepArgs& a, R v, R phi) {
    return a.pf_coupling == 1 ? v : v / fvm_A<R>(a, phi);
}
// Thermal diffusivity of a cell: alpha_l (1 - f_s) + alpha_s f_s, f_s = (1 + phi)/2 in [0, 1]
// (the thermal lattice's alpha(phi)).
template <class R>
__device__ inline R fvm_alpha(const StepArgs& a, R phi) {
    R fs = (R)0.5 * ((R)1 + phi);
    fs = fs < (R)0 ? (R)0 : (fs > (R)1 ? (R)1 : fs);
    return (R)a.alpha_l * ((R)1 - fs) + (R)a.alpha_s * fs;
}

// 19-point stencil of phi (periodic phase faces wrap, others replicate the edge).
template <class R>
__device__ inline void fvm_stencil(const StepArgs& a, const Cell& c, const R* ph, R (&P)[27]) {
    uint32_t px[3] = {c.ox[0], c.ox[1], c.ox[2]};
    uint32_t py[3] = {c.oy[0], c.oy[1], c.oy[2]};
    uint32_t pz[3] = {c.oz[0], c.oz[1], c.oz[2]};
    const uint32_t np = (uint32_t)a.g.npmask & c.bits;
    if (np & XLO) px[0] = px[1];
    if (np & XHI) px[2] = px[1];
    if (np & YLO) py[0] = py[1];
    if (np & YHI) py[2] = py[1];
    if (np & ZLO) pz[0] = pz[1];
    if (np & ZHI) pz[2] = pz[1];
#pragma unroll
    for (int i = 0; i < 19; i++)
        P[(cx(i) + 1) * 9 + (cy(i) + 1) * 3 + (cz(i) + 1)] =
            *slot_at(ph, px[cx(i) + 1] + py[cy(i) + 1] + pz[cz(i) + 1]);
}

template <class R>
__global__ void __launch_bounds__(256) k_fvm_phase(const StepArgs a, const R* __restrict__ phi,
                                                   const R* __restrict__ cc, R* __restrict__ phin) {
    Cell c;
    if (!locate(a, c)) return;
    R P[27];
    fvm_stencil<R>(a, c, phi, P);
    const R p0 = P[13];
    const R U = fvm_u<R>(a, *slot_at(cc, c.id), p0);
    const R M[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    const Aniso<R> an{a.aniso_kind, (R)a.eps4, (R)a.aniso[0], (R)a.aniso[1]};
    R as;
    const R S = phase_source<R>(P, U, an, (R)a.lambda, (R)a.pf_dt, (R)a.pf_wdx2, M, false, as);
    const R lap = P[22] + P[4] + P[16] + P[10] + P[14] + P[12] - (R)6 * p0;
    *slot_at(phin, c.id) = p0 + S + (R)(a.pf_dt * a.pf_wdx2) * lap;
}

// x cells per thread of the marching phase kernel
// Cells per thread of the marching kernels.  At 256^3 FP32, 8 measured fastest: 4 and 16 gave
// 1 % and 2.5 % fewer MLUPS (fewer shared faces against fewer waves to hide the walk).
#ifndef MKG_FVM_XC
#define MKG_FVM_XC 8
#endif
constexpr uint32_t kFvmXC = MKG_FVM_XC;

// Face flux J.n of the face of cell P (27-point layout) on side `side` of axis `ax`: the normal
// difference and the transverse central differences averaged over the two cells, through
// face_flux (the same arithmetic as phase_source).
template <class R>
__device__ inline R fvm_face(const R (&P)[27], int ax, int side, const Aniso<R>& an) {
#define PP(dx, dy, dz) P[((dx) + 1) * 9 + ((dy) + 1) * 3 + ((dz) + 1)]
    const int o0 = ax == 0 ? side : 0, o1 = ax == 1 ? side : 0, o2 = ax == 2 ? side : 0;
    R g[3];
#pragma unroll
    for (int b = 0; b < 3; b++) {
        if (b == ax) {
            g[b] = (R)side * (PP(o0, o1, o2) - PP(0, 0, 0));
        } else {
            const int e0 = b == 0, e1 = b == 1, e2 = b == 2;
            g[b] = (R)0.25 * (PP(e0, e1, e2) - PP(-e0, -e1, -e2) + PP(o0 + e0, o1 + e1, o2 + e2) -
                              PP(o0 - e0, o1 - e1, o2 - e2));
        }
    }
#undef PP
    const R M[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    return face_flux<R>(g[0], g[1], g[2], ax, an, M, false);
}

// Marching form of k_fvm_phase for blocks that hold whole z rows (256 % nz == 0, ny a multiple
// of the rows per block): each thread walks XC cells along x.  A face flux is evaluated once
// for both of its cells where they share a thread or a block: the x face carries over in a
// register, the z face passes through LDS (double-buffered, one barrier per plane); y faces
// and the first x face of a walk are evaluated per cell, as are z faces at a non-periodic z
// boundary.  The phi stencil rolls plane by plane (9 loads per cell instead of 19).  The
// cell part (gradient, anisotropy, source) is phase_source's.
template <class R, int XC>
__global__ void __launch_bounds__(256) k_fvm_phase_march(const StepArgs a, const R* __restrict__ phi,
                                                         const R* __restrict__ cc,
                                                         R* __restrict__ phin) {
    __shared__ R jz[2][256];
    const uint32_t nx = (uint32_t)a.nx, ny = (uint32_t)a.ny, nz = (uint32_t)a.nz;
    const uint32_t tid = threadIdx.x;
    const uint32_t rows = 256u / nz;                 // z rows per block
    const uint32_t ygroups = ny / rows;
    const uint32_t bl = remap_block(a, blockIdx.x, gridDim.x);
    const uint32_t yg = bl % ygroups, xc = bl / ygroups;
    const uint32_t y = yg * rows + tid / nz, z = tid % nz;
    const uint32_t x0 = xc * XC;
    const uint32_t xend = x0 + XC < nx ? x0 + XC : nx;
    const uint32_t np = (uint32_t)a.g.npmask;
    const bool zper = (np & (ZLO | ZHI)) == 0u;
    const uint32_t rb = tid - z;  // first thread of this z row
    // neighbour offsets in y and z (periodic faces wrap, others replicate the edge)
    const uint32_t ylo = y == 0 ? ((np & YLO) ? 0u : ny - 1) : y - 1;
    const uint32_t yhi = y == ny - 1 ? ((np & YHI) ? y : 0u) : y + 1;
    const uint32_t zlo = z == 0 ? ((np & ZLO) ? 0u : nz - 1) : z - 1;
    const uint32_t zhi = z == nz - 1 ? ((np & ZHI) ? z : 0u) : z + 1;
    const uint32_t yo[3] = {ylo * nz, y * nz, yhi * nz}, zo[3] = {zlo, z, zhi};
    const uint32_t pl = ny * nz;
    auto xidx = [&](int xx) -> uint32_t {
        if (xx < 0) return (np & XLO) ? 0u : nx - 1;
        if (xx >= (int)nx) return (np & XHI) ? nx - 1 : 0u;
        return (uint32_t)xx;
    };
    R Q[3][9];  // planes x-1, x, x+1: [(dy+1)*3 + (dz+1)]
    auto load_plane = [&](R (&q)[9], int xx) {
        const uint32_t b = xidx(xx) * pl;
#pragma unroll
        for (int dy = 0; dy < 3; dy++)
#pragma unroll
            for (int dz = 0; dz < 3; dz++) q[dy * 3 + dz] = *slot_at(phi, b + yo[dy] + zo[dz]);
    };
    load_plane(Q[0], (int)x0 - 1);
    load_plane(Q[1], (int)x0);
    const Aniso<R> an{a.aniso_kind, (R)a.eps4, (R)a.aniso[0], (R)a.aniso[1]};
    R jx_carry = 0;
    int buf = 0;
    for (uint32_t x = x0; x < xend; x++) {
        load_plane(Q[2], (int)x + 1);
        R P[27];
#pragma unroll
        for (int k = 0; k < 9; k++) {
            P[k] = Q[0][k];
            P[9 + k] = Q[1][k];
            P[18 + k] = Q[2][k];
        }
        const uint32_t id = x * pl + y * nz + z;
        const R jzp = fvm_face<R>(P, 2, 1, an);
        jz[buf][tid] = jzp;
        const R jxp = fvm_face<R>(P, 0, 1, an);
        const R jxm = x == x0 ? fvm_face<R>(P, 0, -1, an) : jx_carry;
        const R jyp = fvm_face<R>(P, 1, 1, an);
        const R jym = fvm_face<R>(P, 1, -1, an);
        __syncthreads();
        R jzm;
        if (z == 0 && !zper)
            jzm = fvm_face<R>(P, 2, -1, an);
        else
            jzm = jz[buf][rb + (z == 0 ? nz - 1 : z - 1)];
        const R divJ = jxp - jxm + jyp - jym + jzp - jzm;
        const R p0 = P[13];
        const R lap = P[22] + P[4] + P[16] + P[10] + P[14] + P[12] - (R)6 * p0;
        const R U = fvm_u<R>(a, *slot_at(cc, id), p0);
        // cell part of phase_source: LB-weighted cell gradient, anisotropy, source
        R gr0 = 0, gr1 = 0, gr2 = 0;
#pragma unroll
        for (int i = 1; i < 19; i++) {
            const R v = (R)(3.0 * wgt<19>(i)) * P[(cx(i) + 1) * 9 + (cy(i) + 1) * 3 + (cz(i) + 1)];
            if (cx(i)) gr0 += (R)cx(i) * v;
            if (cy(i)) gr1 += (R)cy(i) * v;
            if (cz(i)) gr2 += (R)cz(i) * v;
        }
        R as = 1;
        const R gg = gr0 * gr0 + gr1 * gr1 + gr2 * gr2;
        if (gg > (R)1e-30) {
            R n0, n1, n2, D0, D1, D2;
            unit3<R>(gg, gr0, gr1, gr2, n0, n1, n2);
            as = aniso_eval<R>(an, n0, n1, n2, D0, D1, D2);
        }
        const R d = (R)1 - p0 * p0;
        const R Qs = (p0 - (R)a.lambda * U * d) * d;
        const R a2 = as * as;
        const R S = (R)a.pf_dt / a2 * ((divJ - a2 * lap) * (R)a.pf_wdx2 + Qs);
        *slot_at(phin, id) = p0 + S + (R)(a.pf_dt * a.pf_wdx2) * lap;
        jx_carry = jxp;
        buf ^= 1;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            Q[0][k] = Q[1][k];
            Q[1][k] = Q[2][k];
        }
    }
}

// Heat equation with the latent source on the faces of the solute update (one cell per
// thread): Dirichlet thermal faces fix T at the face (J_out = -2 alpha (T_w - T)), bounce
// faces carry no flux.
template <class R>
__global__ void __launch_bounds__(256) k_fvm_thermal(const StepArgs a, const R* __restrict__ phi,
                                                     const R* __restrict__ phin,
                                                     const R* __restrict__ T,
                                                     R* __restrict__ Tn) {
    Cell c;
    if (!locate(a, c)) return;
    const R p0 = *slot_at(phi, c.id);
    const R T0 = *slot_at(T, c.id);
    const R a0 = fvm_alpha<R>(a, p0);
    const uint32_t nph = (uint32_t)a.h.npmask & c.bits;
    R out = 0;
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
#pragma unroll
        for (int sd = 0; sd < 2; sd++) {
            const int s = sd == 0 ? 1 : -1;
            const int face = 2 * ax + (s > 0 ? 1 : 0);
            if (nph & (1u << face)) {
                if (pick6(face, a.h.bc) == kBcDirichlet)
                    out += (R)-2 * a0 * ((R)pick6(face, a.h.fval) - T0);
                continue;
            }
            const uint32_t nb = nbr(c, ax == 0 ? s : 0, ax == 1 ? s : 0, ax == 2 ? s : 0);
            const R af = (R)0.5 * (a0 + fvm_alpha<R>(a, *slot_at(phi, nb)));
            out += -af * (*slot_at(T, nb) - T0);
        }
    }
    *slot_at(Tn, c.id) = T0 - out + (R)a.latent_T * (*slot_at(phin, c.id) - p0);
}

template <class R>
__global__ void __launch_bounds__(256) k_fvm_solute(const StepArgs a, const R* __restrict__ phi,
                                                    const R* __restrict__ phin,
                                                    const R* __restrict__ cc,
                                                    R* __restrict__ cn) {
    Cell c;
    if (!locate(a, c)) return;
    R P[27];
    fvm_stencil<R>(a, c, phi, P);
#define PP(dx, dy, dz) P[((dx) + 1) * 9 + ((dy) + 1) * 3 + ((dz) + 1)]
    const R p0 = P[13];
    const R U0 = *slot_at(cc, c.id) / fvm_A<R>(a, p0);
    const R q0 = fvm_q<R>(a, p0);
    const R d0 = *slot_at(phin, c.id) - p0;
    const R D = (R)a.D_l, atw = (R)a.at_w, km = (R)1 - (R)a.k_part;
    const uint32_t npc = (uint32_t)a.c.npmask & c.bits;
    R out = 0;
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
#pragma unroll
        for (int sd = 0; sd < 2; sd++) {
            const int s = sd == 0 ? 1 : -1;
            const int face = 2 * ax + (s > 0 ? 1 : 0);
            if (npc & (1u << face)) {
                // non-periodic solute face: Dirichlet fixes U at the face, others carry no flux
                if (pick6(face, a.c.bc) == kBcDirichlet)
                    out += (R)-2 * D * q0 * ((R)pick6(face, a.c.fval) - U0);
                continue;
            }
            const int o0 = ax == 0 ? s : 0, o1 = ax == 1 ? s : 0, o2 = ax == 2 ? s : 0;
            const uint32_t nb = nbr(c, o0, o1, o2);
            const R pn = PP(o0, o1, o2);
            const R Un = *slot_at(cc, nb) / fvm_A<R>(a, pn);
            const R qf = (R)0.5 * (q0 + fvm_q<R>(a, pn));
            R J = -D * qf * (Un - U0);
            if (atw != (R)0) {
                R g[3];
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    if (b == ax) {
                        g[b] = (R)s * (pn - p0);
                    } else {
                        const int e0 = b == 0, e1 = b == 1, e2 = b == 2;
                        g[b] = (R)0.25 * (PP(e0, e1, e2) - PP(-e0, -e1, -e2) +
                                          PP(o0 + e0, o1 + e1, o2 + e2) -
                                          PP(o0 - e0, o1 - e1, o2 - e2));
                    }
                }
                const R gg = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
                if (gg > (R)1e-20) {
                    const R Uf = (R)0.5 * (U0 + Un);
                    const R df = (R)0.5 * (d0 + (*slot_at(phin, nb) - pn));
                    const R ja = -atw * ((R)1 + km * Uf) * df * g[ax] * inv_sqrt<R>(gg);
                    J += (R)s * ja;
                }
            }
            out += J;
        }
    }
#undef PP
    *slot_at(cn, c.id) = *slot_at(cc, c.id) - out + (R)0.5 * d0;
}

// Outward solute flux of a cell through its interior face on side s of axis ax (the
// k_fvm_solute form): diffusion -D q_f (U_n - U_0) and the anti-trapping current.
template <class R>
__device__ inline R sol_face(const StepArgs& a, const R (&P)[27], int ax, int s, R U0, R q0,
                             R d0, R Un, R pn, R dn) {
#define PP(dx, dy, dz) P[((dx) + 1) * 9 + ((dy) + 1) * 3 + ((dz) + 1)]
    const R D = (R)a.D_l, atw = (R)a.at_w, km = (R)1 - (R)a.k_part;
    const R p0 = PP(0, 0, 0);
    const R qf = (R)0.5 * (q0 + fvm_q<R>(a, pn));
    R J = -D * qf * (Un - U0);
    if (atw != (R)0) {
        const int o0 = ax == 0 ? s : 0, o1 = ax == 1 ? s : 0, o2 = ax == 2 ? s : 0;
        R g[3];
#pragma unroll
        for (int b = 0; b < 3; b++) {
            if (b == ax) {
                g[b] = (R)s * (pn - p0);
            } else {
                const int e0 = b == 0, e1 = b == 1, e2 = b == 2;
                g[b] = (R)0.25 * (PP(e0, e1, e2) - PP(-e0, -e1, -e2) + PP(o0 + e0, o1 + e1, o2 + e2) -
                                  PP(o0 - e0, o1 - e1, o2 - e2));
            }
        }
        const R gg = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
        if (gg > (R)1e-20) {
            const R Uf = (R)0.5 * (U0 + Un);
            const R df = (R)0.5 * (d0 + dn);
            J += (R)s * (-atw * ((R)1 + km * Uf) * df * g[ax] * inv_sqrt<R>(gg));
        }
    }
#undef PP
    return J;
}

// Outward flux through a non-periodic solute face: Dirichlet fixes U at the face, the others
// carry none.
template <class R>
__device__ inline R sol_wall(const StepArgs& a, int face, R U0, R q0) {
    if (pick6(face, a.c.bc) == kBcDirichlet)
        return (R)-2 * (R)a.D_l * q0 * ((R)pick6(face, a.c.fval) - U0);
    return (R)0;
}

// Marching form of k_fvm_solute (same block shape as k_fvm_phase_march): the outward flux of
// a face is the negative of its neighbour's, so each thread evaluates its high x, y, z faces,
// takes the low x face from its previous cell and the low z face from LDS; y faces, the first
// low x face of a walk and the faces on non-periodic boundaries are evaluated per cell.
template <class R, int XC>
__global__ void __launch_bounds__(256) k_fvm_solute_march(const StepArgs a, const R* __restrict__ phi,
                                                          const R* __restrict__ phin,
                                                          const R* __restrict__ cc,
                                                          R* __restrict__ cn) {
    __shared__ R jz[2][256];
    const uint32_t nx = (uint32_t)a.nx, ny = (uint32_t)a.ny, nz = (uint32_t)a.nz;
    const uint32_t tid = threadIdx.x;
    const uint32_t rows = 256u / nz;
    const uint32_t ygroups = ny / rows;
    const uint32_t bl = remap_block(a, blockIdx.x, gridDim.x);
    const uint32_t yg = bl % ygroups, xc = bl / ygroups;
    const uint32_t y = yg * rows + tid / nz, z = tid % nz;
    const uint32_t x0 = xc * XC;
    const uint32_t xend = x0 + XC < nx ? x0 + XC : nx;
    const uint32_t npg = (uint32_t)a.g.npmask, npc = (uint32_t)a.c.npmask;
    const uint32_t rb = tid - z;
    // phi stencil offsets (phase faces: periodic wrap, others replicate the edge)
    const uint32_t ylo = y == 0 ? ((npg & YLO) ? 0u : ny - 1) : y - 1;
    const uint32_t yhi = y == ny - 1 ? ((npg & YHI) ? y : 0u) : y + 1;
    const uint32_t zlo = z == 0 ? ((npg & ZLO) ? 0u : nz - 1) : z - 1;
    const uint32_t zhi = z == nz - 1 ? ((npg & ZHI) ? z : 0u) : z + 1;
    const uint32_t yo[3] = {ylo * nz, y * nz, yhi * nz}, zo[3] = {zlo, z, zhi};
    const uint32_t pl = ny * nz;
    auto xidx = [&](int xx) -> uint32_t {
        if (xx < 0) return (npg & XLO) ? 0u : nx - 1;
        if (xx >= (int)nx) return (npg & XHI) ? nx - 1 : 0u;
        return (uint32_t)xx;
    };
    // solute neighbours (periodic axes are the same for both fields)
    const uint32_t yw_lo = y == 0 ? ny - 1 : y - 1, yw_hi = y == ny - 1 ? 0u : y + 1;
    const uint32_t zw_hi = z == nz - 1 ? 0u : z + 1;
    const bool ylo_wall = y == 0 && (npc & YLO), yhi_wall = y == ny - 1 && (npc & YHI);
    const bool zlo_wall = z == 0 && (npc & ZLO), zhi_wall = z == nz - 1 && (npc & ZHI);
    R Q[3][9];
    auto load_plane = [&](R (&q)[9], int xx) {
        const uint32_t b = xidx(xx) * pl;
#pragma unroll
        for (int dy = 0; dy < 3; dy++)
#pragma unroll
            for (int dz = 0; dz < 3; dz++) q[dy * 3 + dz] = *slot_at(phi, b + yo[dy] + zo[dz]);
    };
    load_plane(Q[0], (int)x0 - 1);
    load_plane(Q[1], (int)x0);
    const uint32_t col = y * nz + z;
    R c0 = *slot_at(cc, x0 * pl + col), pn0 = *slot_at(phin, x0 * pl + col);
    R jx_carry = 0;
    int buf = 0;
    for (uint32_t x = x0; x < xend; x++) {
        load_plane(Q[2], (int)x + 1);
        R P[27];
#pragma unroll
        for (int k = 0; k < 9; k++) {
            P[k] = Q[0][k];
            P[9 + k] = Q[1][k];
            P[18 + k] = Q[2][k];
        }
        const uint32_t id = x * pl + col;
        const R p0 = P[13];
        const R U0 = c0 / fvm_A<R>(a, p0);
        const R q0 = fvm_q<R>(a, p0);
        const R d0 = pn0 - p0;
        auto interior = [&](int ax, int s, uint32_t nb, R pn) {
            const R Un = *slot_at(cc, nb) / fvm_A<R>(a, pn);
            return sol_face<R>(a, P, ax, s, U0, q0, d0, Un, pn, *slot_at(phin, nb) - pn);
        };
        // x+ face (and the next cell's values)
        const bool xhi_wall = x == nx - 1 && (npc & XHI);
        const uint32_t xn = x == nx - 1 ? 0u : x + 1;
        const R c1 = *slot_at(cc, xn * pl + col), pn1 = *slot_at(phin, xn * pl + col);
        R jxp;
        if (xhi_wall) {
            jxp = sol_wall<R>(a, 1, U0, q0);
        } else {
            const R pn = P[22];
            const R Un = c1 / fvm_A<R>(a, pn);
            jxp = sol_face<R>(a, P, 0, 1, U0, q0, d0, Un, pn, pn1 - pn);
        }
        // z+ face through LDS (the z neighbour's low face)
        const R jzp = zhi_wall ? sol_wall<R>(a, 5, U0, q0)
                               : interior(2, 1, x * pl + y * nz + zw_hi, P[14]);
        jz[buf][tid] = jzp;
        // x- face: the previous cell's x+ face, or evaluated at the start of the walk
        R jxm;
        if (x != x0) {
            jxm = -jx_carry;
        } else if (x == 0 && (npc & XLO)) {
            jxm = sol_wall<R>(a, 0, U0, q0);
        } else {
            jxm = interior(0, -1, (x == 0 ? nx - 1 : x - 1) * pl + col, P[4]);
        }
        const R jyp = yhi_wall ? sol_wall<R>(a, 3, U0, q0) : interior(1, 1, x * pl + yw_hi * nz + z, P[16]);
        const R jym = ylo_wall ? sol_wall<R>(a, 2, U0, q0) : interior(1, -1, x * pl + yw_lo * nz + z, P[10]);
        __syncthreads();
        R jzm;
        if (zlo_wall)
            jzm = sol_wall<R>(a, 4, U0, q0);
        else
            jzm = -jz[buf][rb + (z == 0 ? nz - 1 : z - 1)];
        const R out = jxp + jxm + jyp + jym + jzp + jzm;
        *slot_at(cn, id) = c0 - out + (R)0.5 * d0;
        jx_carry = jxp;
        c0 = c1;
        pn0 = pn1;
        buf ^= 1;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            Q[0][k] = Q[1][k];
            Q[1][k] = Q[2][k];
        }
    }
}

}  // namespace mkg

using namespace mkg;

struct mkg_fvm_ctx {
    mkg_params p;
    int fp32 = 0;
    size_t esz = 8;
    uint32_t N = 0, S = 0, nblk = 0;
    hipStream_t stream = nullptr;
    void* phi[2] = {nullptr, nullptr};
    void* c[2] = {nullptr, nullptr};  // c = A(phi) U, or T when thermal
    int thermal = 0;  // the potential is the temperature (evolution_thermal)
    int cur = 0;  // buffer of the current level
    uint32_t march_blk = 0;   // blocks of the marching kernels (0: the per-cell kernels)
    int xcd_chunk = 0;        // block order of the launches, fixed at creation
    int8_t* flag = nullptr;   // all fluid: the reductions read flags
    double* scratch = nullptr;
    double* partials = nullptr;
    double* red_out = nullptr;
};

namespace {

int ffail(const std::string& msg) {
    set_error(msg);
    return -1;
}

#define FCHK(expr)                                                                     \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return ffail(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

StepArgs fvm_args(const mkg_fvm_ctx* h) {
    StepArgs a;
    std::memset(&a, 0, sizeof(a));
    const mkg_params& p = h->p;
    a.nx = p.nx;
    a.ny = p.ny;
    a.nz = p.nz;
    a.N = h->N;
    a.S = h->S;
    a.nblk = h->nblk;
    a.flag = h->flag;
    a.xcd_chunk = h->xcd_chunk;
    int npg = 0, npc = 0;
    for (int f = 0; f < 6; f++) {
        a.g.bc[f] = p.bc_g[f];
        a.c.bc[f] = p.bc_c[f];
        a.c.fval[f] = p.C_face[f];
        if (p.bc_g[f] != MKG_BC_PERIODIC) npg |= 1 << f;
        if (p.bc_c[f] != MKG_BC_PERIODIC) npc |= 1 << f;
    }
    a.g.npmask = npg;
    a.c.npmask = npc;
    int nph = 0;
    for (int f = 0; f < 6; f++) {
        a.h.bc[f] = p.bc_h[f];
        a.h.fval[f] = p.T_face[f];
        if (p.bc_h[f] != MKG_BC_PERIODIC) nph |= 1 << f;
    }
    a.h.npmask = nph;
    a.alpha_l = p.alpha_l;
    a.alpha_s = p.alpha_s;
    a.latent_T = p.latent_T;
    a.pf_coupling = h->thermal ? 1 : 2;
    a.D_l = p.D_l;
    a.D_s = p.D_s;
    a.k_part = p.k_part;
    a.q_min = p.q_min;
    a.at_w = p.at_coef * std::sqrt(p.pf_wdx2 > 0 ? p.pf_wdx2 : 0.0);
    a.pf_dt = p.pf_dt;
    a.pf_wdx2 = p.pf_wdx2;
    a.eps4 = p.eps4;
    a.lambda = p.lambda;
    a.aniso_kind = p.aniso_kind;
    a.aniso[0] = p.aniso[0];
    a.aniso[1] = p.aniso[1];
    return a;
}

void fvm_destroy_impl(mkg_fvm_ctx* h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void* bufs[] = {h->phi[0], h->phi[1], h->c[0], h->c[1], h->flag, h->scratch, h->partials,
                    h->red_out};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

}  // namespace

extern "C" {

int mkg_fvm_create(const mkg_params* p, mkg_fvm_handle* out) {
    if (!out) return ffail("out is null");
    *out = nullptr;
    if (!p) return ffail("params is null");
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) return ffail("grid dimensions must be >= 1");
    if (p->precision != MKG_PREC_FP64 && p->precision != MKG_PREC_FP32)
        return ffail("precision must be MKG_PREC_FP64 or MKG_PREC_FP32");
    if (p->ngrains != 0) return ffail("the finite-volume solver has no grain rotations (ngrains 0)");
    if (p->aniso_kind < MKG_ANISO_CUBIC || p->aniso_kind > MKG_ANISO_FACET)
        return ffail("aniso_kind out of range");
    if (p->aniso_kind == MKG_ANISO_FACET && !(p->aniso[1] > 0))
        return ffail("the facet anisotropy needs a rounding nu > 0");
    if (!(p->k_part > 0) || !(p->q_min > 0)) return ffail("k_part and q_min must be > 0");
    // the potential: the temperature (thermal = 1, evolution_thermal) or the solute
    const int32_t* bcp = p->thermal ? p->bc_h : p->bc_c;
    for (int f = 0; f < 6; f++) {
        if (p->bc_g[f] != MKG_BC_PERIODIC && p->bc_g[f] != MKG_BC_BOUNCE)
            return ffail("finite-volume phase faces are periodic or bounce (zero flux)");
        if (bcp[f] != MKG_BC_PERIODIC && bcp[f] != MKG_BC_BOUNCE && bcp[f] != MKG_BC_DIRICHLET)
            return ffail("finite-volume solute / thermal faces are periodic, bounce (zero flux) "
                         "or Dirichlet");
    }
    for (int a = 0; a < 3; a++) {
        const bool pg = p->bc_g[2 * a] == 0, pc = bcp[2 * a] == 0;
        if (pg != (p->bc_g[2 * a + 1] == 0) || pc != (bcp[2 * a + 1] == 0))
            return ffail("a periodic axis needs periodic codes on both faces");
        if (pg != pc)
            return ffail("the phase and solute / thermal fields must be periodic on the same axes");
    }
    if (p->thermal && (!(p->alpha_l > 0) || !(p->alpha_s > 0) || p->alpha_l > 1.0 / 6.0 ||
                       p->alpha_s > 1.0 / 6.0))
        return ffail("the explicit heat equation needs 0 < alpha_l, alpha_s <= 1/6");
    const double n = (double)p->nx * p->ny * p->nz;
    const double esz = p->precision == MKG_PREC_FP32 ? 4.0 : 8.0;
    if ((n + 65536.0) * esz >= 4294967296.0) return ffail("grid too large for 32-bit offsets");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return ffail("no HIP device available");
    mkg_fvm_ctx* h = new mkg_fvm_ctx();
    h->p = *p;
    h->thermal = p->thermal ? 1 : 0;
    h->fp32 = p->precision == MKG_PREC_FP32;
    h->esz = h->fp32 ? 4 : 8;
    h->N = (uint32_t)n;
    h->S = ((h->N + 127u) & ~127u) + 1024u;
    h->nblk = (h->N + 255u) / 256u;
    h->xcd_chunk = xcd_chunk_setting(h->esz);
    // the marching kernels need whole z rows per block
    if (256u % (uint32_t)p->nz == 0u && (uint32_t)p->ny % (256u / (uint32_t)p->nz) == 0u) {
        const uint32_t rows = 256u / (uint32_t)p->nz;
        h->march_blk = (((uint32_t)p->nx + kFvmXC - 1) / kFvmXC) * ((uint32_t)p->ny / rows);
    }
    if (const char* env = std::getenv("MKG_FVM_MARCH"))
        if (std::atoi(env) == 0) h->march_blk = 0;
    auto bad = [&](hipError_t e, const char* what) {
        fvm_destroy_impl(h);
        return ffail(std::string(what) + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipStreamCreate(&h->stream)) != hipSuccess) return bad(e, "hipStreamCreate");
    for (int b = 0; b < 2; b++) {
        if ((e = hipMalloc(&h->phi[b], (size_t)h->S * h->esz)) != hipSuccess) return bad(e, "hipMalloc");
        if ((e = hipMalloc(&h->c[b], (size_t)h->S * h->esz)) != hipSuccess) return bad(e, "hipMalloc");
    }
    if ((e = hipMalloc((void**)&h->flag, h->S)) != hipSuccess) return bad(e, "hipMalloc");
    if ((e = hipMemsetAsync(h->flag, 0, h->S, h->stream)) != hipSuccess) return bad(e, "hipMemsetAsync");
    if ((e = hipMalloc((void**)&h->scratch, (size_t)h->N * sizeof(double))) != hipSuccess)
        return bad(e, "hipMalloc");
    if ((e = hipMalloc((void**)&h->partials, (size_t)(h->nblk + 1) * 4 * sizeof(double))) != hipSuccess)
        return bad(e, "hipMalloc");
    if ((e = hipMalloc((void**)&h->red_out, 8 * sizeof(double))) != hipSuccess) return bad(e, "hipMalloc");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return bad(e, "hipStreamSynchronize");
    *out = h;
    return 0;
}

int mkg_fvm_destroy(mkg_fvm_handle h) {
    fvm_destroy_impl(h);
    return 0;
}

int mkg_fvm_set(mkg_fvm_handle h, int which, const double* host, size_t count) {
    if (!h) return ffail("null handle");
    if (!host) return ffail("host buffer is null");
    if (which != MKG_FVM_PHI && which != (h->thermal ? MKG_FVM_TEMP : MKG_FVM_CONC))
        return ffail(h->thermal ? "set takes MKG_FVM_PHI or MKG_FVM_TEMP (thermal handle)"
                                : "set takes MKG_FVM_PHI or MKG_FVM_CONC");
    if (count != h->N) return ffail("count does not match the field size");
    void* dst = which == MKG_FVM_PHI ? h->phi[h->cur] : h->c[h->cur];
    FCHK(hipMemcpyAsync(h->scratch, host, h->N * sizeof(double), hipMemcpyHostToDevice, h->stream));
    launch_from_double(h->fp32, h->scratch, dst, h->N, h->stream);
    FCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int mkg_fvm_step(mkg_fvm_handle h, int nsteps, int evolve_phase) {
    if (!h) return ffail("null handle");
    if (nsteps < 0) return ffail("nsteps must be >= 0");
    const StepArgs a = fvm_args(h);
    dim3 grid(h->nblk), block(256);
    for (int s = 0; s < nsteps; s++) {
        const int o = h->cur, n = h->cur ^ 1;
        // with the phase held fixed both kernels see phi' = phi
        void* phin = evolve_phase ? h->phi[n] : h->phi[o];
        if (h->fp32) {
            if (evolve_phase && h->march_blk)
                k_fvm_phase_march<float, kFvmXC><<<h->march_blk, 256, 0, h->stream>>>(
                    a, (const float*)h->phi[o], (const float*)h->c[o], (float*)phin);
            else if (evolve_phase)
                k_fvm_phase<float><<<grid, block, 0, h->stream>>>(
                    a, (const float*)h->phi[o], (const float*)h->c[o], (float*)phin);
            if (h->thermal)
                k_fvm_thermal<float><<<grid, block, 0, h->stream>>>(
                    a, (const float*)h->phi[o], (const float*)phin, (const float*)h->c[o],
                    (float*)h->c[n]);
            else if (h->march_blk)
                k_fvm_solute_march<float, kFvmXC><<<h->march_blk, 256, 0, h->stream>>>(
                    a, (const float*)h->phi[o], (const float*)phin, (const float*)h->c[o],
                    (float*)h->c[n]);
            else
                k_fvm_solute<float><<<grid, block, 0, h->stream>>>(
                    a, (const float*)h->phi[o], (const float*)phin, (const float*)h->c[o],
                    (float*)h->c[n]);
        } else {
            if (evolve_phase && h->march_blk)
                k_fvm_phase_march<double, kFvmXC><<<h->march_blk, 256, 0, h->stream>>>(
                    a, (const double*)h->phi[o], (const double*)h->c[o], (double*)phin);
            else if (evolve_phase)
                k_fvm_phase<double><<<grid, block, 0, h->stream>>>(
                    a, (const double*)h->phi[o], (const double*)h->c[o], (double*)phin);
            if (h->thermal)
                k_fvm_thermal<double><<<grid, block, 0, h->stream>>>(
                    a, (const double*)h->phi[o], (const double*)phin, (const double*)h->c[o],
                    (double*)h->c[n]);
            else if (h->march_blk)
                k_fvm_solute_march<double, kFvmXC><<<h->march_blk, 256, 0, h->stream>>>(
                    a, (const double*)h->phi[o], (const double*)phin, (const double*)h->c[o],
                    (double*)h->c[n]);
            else
                k_fvm_solute<double><<<grid, block, 0, h->stream>>>(
                    a, (const double*)h->phi[o], (const double*)phin, (const double*)h->c[o],
                    (double*)h->c[n]);
        }
        if (!evolve_phase)
            FCHK(hipMemcpyAsync(h->phi[n], h->phi[o], (size_t)h->N * h->esz,
                                hipMemcpyDeviceToDevice, h->stream));
        h->cur = n;
    }
    FCHK(hipGetLastError());
    return 0;
}

int mkg_fvm_sync(mkg_fvm_handle h) {
    if (!h) return ffail("null handle");
    FCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int mkg_fvm_get(mkg_fvm_handle h, int which, double* host, size_t count) {
    if (!h) return ffail("null handle");
    if (!host) return ffail("host buffer is null");
    if (count != h->N) return ffail("count does not match the field size");
    const int pot = h->thermal ? MKG_FVM_TEMP : MKG_FVM_CONC;
    if (which == MKG_FVM_PHI || which == pot) {
        const void* src = which == MKG_FVM_PHI ? h->phi[h->cur] : h->c[h->cur];
        launch_to_double(h->fp32, src, h->scratch, h->N, h->stream);
        FCHK(hipMemcpyAsync(host, h->scratch, h->N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        FCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    if (which == MKG_FVM_U && !h->thermal) {
        if (mkg_fvm_get(h, MKG_FVM_CONC, host, count)) return -1;
        double* ph = new double[count];
        const int rc = mkg_fvm_get(h, MKG_FVM_PHI, ph, count);
        if (rc == 0) {
            const double k = h->p.k_part;
            for (size_t i = 0; i < count; i++) host[i] /= 0.5 * ((1 + k) - (1 - k) * ph[i]);
        }
        delete[] ph;
        return rc;
    }
    return ffail(h->thermal ? "a thermal handle has MKG_FVM_PHI and MKG_FVM_TEMP"
                            : "a solute handle has MKG_FVM_PHI, MKG_FVM_CONC and MKG_FVM_U");
}

int mkg_fvm_sums(mkg_fvm_handle h, double* out) {
    if (!h) return ffail("null handle");
    if (!out) return ffail("out is null");
    const StepArgs a = fvm_args(h);
    for (int k = 0; k < 2; k++) {
        launch_to_double(h->fp32, k == 0 ? h->phi[h->cur] : h->c[h->cur], h->scratch, h->N,
                         h->stream);
        int nout = 0;
        launch_reduce(4, a, nullptr, nullptr, nullptr, nullptr, h->scratch, h->partials,
                      h->red_out, &nout, h->stream);
        FCHK(hipMemcpyAsync(out + k, h->red_out, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    }
    FCHK(hipStreamSynchronize(h->stream));
    return 0;
}

}  // extern "C"
