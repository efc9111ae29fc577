// Discrete adjoint of the fused phase-field (D3Q19) + thermal (D3Q7) step, the forward model
// of the inverse problem (no flow lattice).  Hand-written reverse of k_phase_moments + k_step:
//
//   forward step n at cell x (canonical pull view):
//     phi = sum_i g_i,  T = sum_k h_k,  S = S(phi over the 19-point stencil, T; eps4)
//     g*_i = g_i - w_g (g_i - w_i phi) + w_i S
//     w_h  = 1 / (1/2 + kappa alpha(phi) / cs^2)
//     h*_k = h_k - w_h (h_k - w_k T) + w_k L (phi - prev)
//     g'_i(x + c_i) = g*_i(x), h'_k(x + c_k) = h*_k(x)  (face rules), prev' = phi
//
// The reverse of one step is two launches:
//   k_adj_collide  gathers the streamed adjoints, reverses both collisions and the source
//                  (reverse mode through the face fluxes and the cell anisotropy, with the
//                  exact Hessian of the Karma-Rappel energy), writes the adjoint populations
//                  and the stencil contributions, and accumulates dL/dkappa and the FP64 block
//                  partials of dL/deps4;
//   k_adj_qsum /   the transpose of the stencil: the part of the phi adjoint that every
//   k_adj_stencil  population of the cell receives, folded in when the next (earlier) step
//                  gathers.  When a block holds whole z rows the contributions leave
//                  k_adj_collide as nine column sums per cell (36 B in FP32) and k_adj_qsum
//                  adds them; otherwise as 19 per-direction values (76 B) that k_adj_stencil
//                  gathers with the face clamps.
// Adjoint populations live in plain two-lattice canonical buffers (Q x N).
#include "device.h"
#include "fadj.h"

namespace mkg {

// ---------------------------------------------------------------- source derivatives
// Face flux and its derivatives in the crystal frame: with n = G/|G|, s4 = sum n^4,
//   a = 1 - 3 e + 4 e s4,  da_k = |G| da/dG_k = 16 e (n_k^3 - s4 n_k),
//   J_k = |G| (a^2 n_k + a da_k)                         (= dE/dG_k, E = |G|^2 a^2 / 2)
//   H_kl = dJ_k/dG_l = a^2 d_kl + 2a (n_k da_l + n_l da_k) + da_k da_l + a dda_kl,
//   dda_kl = |G|^2 d2a/dG_k dG_l = 16 e [3 n_k^2 d_kl - 4 n_k^3 n_l - 4 n_l^3 n_k - s4 d_kl
//                                         + 6 s4 n_k n_l]
//   dJ_k/de = |G| (2 a ae n_k + ae da_k + a da_k / e),  ae = -3 + 4 s4.
// Only the product H v with v = the lab axis in crystal coordinates is needed; it is formed
// from the dot products n.v, da.v and n^3.v (about a third of the work of the full matrix):
//   (H v)_k = a^2 v_k + 2a (n_k da.v + da_k n.v) + da_k da.v
//             + 16 e a (3 n_k^2 v_k - 4 n_k^3 n.v - 4 n_k n^3.v - s4 v_k + 6 s4 n_k n.v).
// Returns J_axis (lab) and fills dJ/dg (lab, 3) and dJ/de.
template <class R>
__device__ inline R face_flux_d(const R (&g)[3], int axis, R eps4, const R (&M)[9], bool rot,
                                R (&dJdg)[3], R& dJde) {
    R G[3];
    if (rot) {
        for (int k = 0; k < 3; k++) G[k] = M[0 * 3 + k] * g[0] + M[1 * 3 + k] * g[1] + M[2 * 3 + k] * g[2];
    } else {
        G[0] = g[0];
        G[1] = g[1];
        G[2] = g[2];
    }
    const R r2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
    if (!(r2 > (R)1e-30)) {
        dJdg[0] = dJdg[1] = dJdg[2] = 0;
        dJdg[axis] = 1;
        dJde = 0;
        return g[axis];
    }
    const R r = sqrt(r2);
    const R ir = (R)1 / r;
    const R n[3] = {G[0] * ir, G[1] * ir, G[2] * ir};
    const R n3[3] = {n[0] * n[0] * n[0], n[1] * n[1] * n[1], n[2] * n[2] * n[2]};
    const R s4 = n3[0] * n[0] + n3[1] * n[1] + n3[2] * n[2];
    const R a = (R)1 - (R)3 * eps4 + (R)4 * eps4 * s4;
    const R ae = (R)-3 + (R)4 * s4;
    // v = row `axis` of M (lab axis in crystal coordinates)
    R v[3];
    for (int k = 0; k < 3; k++) v[k] = rot ? M[axis * 3 + k] : (R)(k == axis);
    R da1[3], da[3];
    R nv = 0, dav = 0, n3v = 0, Jl = 0, Jel = 0;
    for (int k = 0; k < 3; k++) {
        da1[k] = (R)16 * (n3[k] - s4 * n[k]);
        da[k] = eps4 * da1[k];
        nv += n[k] * v[k];
        dav += da[k] * v[k];
        n3v += n3[k] * v[k];
        Jl += v[k] * (a * a * n[k] + a * da[k]);
        Jel += v[k] * ((R)2 * a * ae * n[k] + ae * da[k] + a * da1[k]);
    }
    const R a2 = a * a, c16 = (R)16 * eps4 * a;
    R w[3];
    for (int k = 0; k < 3; k++)
        w[k] = a2 * v[k] + (R)2 * a * (n[k] * dav + da[k] * nv) + da[k] * dav +
               c16 * ((R)3 * n[k] * n[k] * v[k] - (R)4 * n3[k] * nv - (R)4 * n[k] * n3v -
                      s4 * v[k] + (R)6 * s4 * n[k] * nv);
    if (rot) {
        for (int b = 0; b < 3; b++) dJdg[b] = M[b * 3 + 0] * w[0] + M[b * 3 + 1] * w[1] + M[b * 3 + 2] * w[2];
    } else {
        dJdg[0] = w[0];
        dJdg[1] = w[1];
        dJdg[2] = w[2];
    }
    dJde = r * Jel;
    return r * Jl;
}

// Cell anisotropy a(grad) with da/dgrad (lab) and da/de.
template <class R>
__device__ inline R cell_as_d(const R (&gr)[3], R eps4, const R (&M)[9], bool rot, R (&dadg)[3],
                              R& dade) {
    R G[3];
    if (rot) {
        for (int k = 0; k < 3; k++) G[k] = M[0 * 3 + k] * gr[0] + M[1 * 3 + k] * gr[1] + M[2 * 3 + k] * gr[2];
    } else {
        G[0] = gr[0];
        G[1] = gr[1];
        G[2] = gr[2];
    }
    const R gg = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
    if (!(gg > (R)1e-30)) {
        dadg[0] = dadg[1] = dadg[2] = 0;
        dade = 0;
        return 1;
    }
    const R r = sqrt(gg);
    R n[3] = {G[0] / r, G[1] / r, G[2] / r};
    const R s4 = pow4(n[0]) + pow4(n[1]) + pow4(n[2]);
    R dc[3];
    for (int k = 0; k < 3; k++) dc[k] = (R)16 * eps4 * (n[k] * n[k] * n[k] - s4 * n[k]) / r;
    if (rot) {
        for (int b = 0; b < 3; b++) dadg[b] = M[b * 3 + 0] * dc[0] + M[b * 3 + 1] * dc[1] + M[b * 3 + 2] * dc[2];
    } else {
        dadg[0] = dc[0];
        dadg[1] = dc[1];
        dadg[2] = dc[2];
    }
    dade = (R)-3 + (R)4 * s4;
    return (R)1 - (R)3 * eps4 + (R)4 * eps4 * s4;
}

#define PIX(dx, dy, dz) (((dx) + 1) * 9 + ((dy) + 1) * 3 + ((dz) + 1))

// Reverse of phase_source: given Sh = dL/dS, accumulates dL/dP (27, corners unused) and
// returns dL/du; adds dL/deps4 to *eh.  The cell anisotropy a comes first (the face terms of
// the reverse scale with 1/a^2), then one pass over the six faces forms each face flux with its
// derivatives and scatters its adjoint at once; the term through a, which needs div J, closes.
template <class R>
__device__ inline R phase_source_adj(const R (&P)[27], R upf, R eps4, R lambda, R pf_dt,
                                     R pf_wdx2, const R (&M)[9], bool rot, R Sh, R (&Ph)[27],
                                     R& eh) {
    R gr[3] = {0, 0, 0};
#pragma unroll
    for (int i = 1; i < 19; i++) {
        const R v = (R)(3.0 * wgt<19>(i)) * P[PIX(cx(i), cy(i), cz(i))];
        if (cx(i)) gr[0] += (R)cx(i) * v;
        if (cy(i)) gr[1] += (R)cy(i) * v;
        if (cz(i)) gr[2] += (R)cz(i) * v;
    }
    R dadg[3], dade;
    const R a = cell_as_d<R>(gr, eps4, M, rot, dadg, dade);
    const R a2 = a * a;
    const R Jh = Sh * pf_dt * pf_wdx2 / a2;          // dL/d(divJ)
    const R lh = -Sh * pf_dt * pf_wdx2;              // dL/d(lap)
    R divJ = 0, e_acc = 0;
#pragma unroll
    for (int k = 0; k < 27; k++) Ph[k] = 0;
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
#pragma unroll
        for (int sd = 0; sd < 2; sd++) {
            const int side = sd == 0 ? 1 : -1;
            const int o0 = ax == 0 ? side : 0, o1 = ax == 1 ? side : 0, o2 = ax == 2 ? side : 0;
            R gf[3];
#pragma unroll
            for (int b = 0; b < 3; b++) {
                if (b == ax) {
                    gf[b] = (R)side * (P[PIX(o0, o1, o2)] - P[PIX(0, 0, 0)]);
                } else {
                    const int e0 = b == 0, e1 = b == 1, e2 = b == 2;
                    gf[b] = (R)0.25 * (P[PIX(e0, e1, e2)] - P[PIX(-e0, -e1, -e2)] +
                                       P[PIX(o0 + e0, o1 + e1, o2 + e2)] -
                                       P[PIX(o0 - e0, o1 - e1, o2 - e2)]);
                }
            }
            R dJ[3], dJe;
            const R J = face_flux_d<R>(gf, ax, eps4, M, rot, dJ, dJe);
            divJ += (R)side * J;
            const R Jf = (R)side * Jh;
            e_acc += Jf * dJe;
            R gh[3];
#pragma unroll
            for (int b = 0; b < 3; b++) gh[b] = Jf * dJ[b];
            gh[ax] += (R)side * lh;
#pragma unroll
            for (int b = 0; b < 3; b++) {
                if (b == ax) {
                    Ph[PIX(o0, o1, o2)] += (R)side * gh[b];
                    Ph[PIX(0, 0, 0)] -= (R)side * gh[b];
                } else {
                    const int e0 = b == 0, e1 = b == 1, e2 = b == 2;
                    const R q = (R)0.25 * gh[b];
                    Ph[PIX(e0, e1, e2)] += q;
                    Ph[PIX(-e0, -e1, -e2)] -= q;
                    Ph[PIX(o0 + e0, o1 + e1, o2 + e2)] += q;
                    Ph[PIX(o0 - e0, o1 - e1, o2 - e2)] -= q;
                }
            }
        }
    }
    const R ph = P[PIX(0, 0, 0)];
    const R d = (R)1 - ph * ph;
    const R Qs = (ph - lambda * upf * d) * d;
    const R Qh = Sh * pf_dt / a2;                    // dL/dQ
    const R ah = Sh * (R)-2 * pf_dt * (pf_wdx2 * divJ + Qs) / (a2 * a);
    e_acc += ah * dade;
    const R grh0 = ah * dadg[0], grh1 = ah * dadg[1], grh2 = ah * dadg[2];
#pragma unroll
    for (int i = 1; i < 19; i++) {
        const R c = (R)(3.0 * wgt<19>(i));
        Ph[PIX(cx(i), cy(i), cz(i))] += c * ((R)cx(i) * grh0 + (R)cy(i) * grh1 + (R)cz(i) * grh2);
    }
    // Q = (ph - lambda u d) d, d = 1 - ph^2
    const R dQdph = d - (R)2 * ph * (ph - lambda * upf * d) + (R)2 * lambda * upf * ph * d;
    Ph[PIX(0, 0, 0)] += Qh * dQdph;
    eh += e_acc;
    return Qh * (-lambda * d * d);
}

// ---------------------------------------------------------------- adjoint streaming
// dL/d(post_i(y)) from the adjoint of the next level's pre-collision populations (canonical
// layout, adj_j(z) = A[j*N + z] + cellwide[z]); reverses the pull rules of the forward step.
// FACES false: the cell touches no non-periodic face of the lattice, so every population was
// pulled by its regular neighbour.
template <int Q, class R, bool FACES = true>
__device__ inline void adj_gather(R* out, const R* __restrict__ A, const R* __restrict__ cw,
                                  const LatArgs& L, const Cell& c, const StepArgs& a, int kind,
                                  uint32_t N) {
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const int dx = cx(i), dy = cy(i), dz = cz(i);
        R v = 0;
        // (a)/(b): where post_i(y) is pulled to
        uint32_t hit = FACES ? c.bits & faces_of(dx, dy, dz) & (uint32_t)L.npmask : 0u;
        if (hit == 0u) {
            const uint32_t t = nbr(c, dx, dy, dz);
            v = ldp(slot_at(A + (size_t)i * N, t)) + (cw ? *slot_at(cw, t) : (R)0);
        } else {
            const int face = __builtin_ctz(hit);
            const int tb = pick6(face, L.bc);
            const R own = *slot_at(A + (size_t)inv(i) * N, c.id) + (cw ? *slot_at(cw, c.id) : (R)0);
            if (tb == kBcOutflow) v = 0;
            else if (tb == kBcDirichlet && kind == kKindScalar) v = -own;
            else v = own;
        }
        // (c): zero-gradient outflow copies post_i(y) into new_i(y)
        uint32_t hin = FACES ? c.bits & faces_of(-dx, -dy, -dz) & (uint32_t)L.npmask : 0u;
        if (hin != 0u) {
            const int face = __builtin_ctz(hin);
            if (pick6(face, L.bc) == kBcOutflow)
                v += *slot_at(A + (size_t)i * N, c.id) + (cw ? *slot_at(cw, c.id) : (R)0);
        }
        out[i] = v;
    }
}

// ---------------------------------------------------------------- grouped stencil transpose
// Ph[d] (d a 19-point stencil offset) is the adjoint of phi read by cell x at
// clampwrap(x + d): a non-periodic face of the phase lattice replicates the cell's own value.
// The contributions are first moved to the offsets they were actually read from (a clamped
// axis folds d onto 0), then summed per (dx, dy) column: the dz = +-1 parts go to the cells at
// z -+ 1 of the same row through LDS (a non-periodic z face has folded them away, so the
// wrap-around reads zero there), and the column sum lands at the target row (x+dx, y+dy) in
// array Q[(dx+1)*3 + (dy+1)].  Every (array, cell) is written by exactly one source cell per
// step (the one at cell - (dx, dy), with periodic wrap), so no array needs clearing; the phi
// adjoint of a cell is the sum of its nine array values (k_adj_qsum).  Every thread of the
// block calls this (one barrier); act: the thread holds a live cell.
template <class R>
__device__ inline void group_stencil_adjoint(const StepArgs& a, const Cell& c, bool act,
                                             R (&Ph)[27], R (*zq)[256], R* __restrict__ Q) {
    const uint32_t np = (uint32_t)a.g.npmask & c.bits;
    if (np & (XLO | XHI)) {
#pragma unroll
        for (int dy = -1; dy <= 1; dy++)
#pragma unroll
            for (int dz = -1; dz <= 1; dz++) {
                if (np & XLO) {
                    Ph[PIX(0, dy, dz)] += Ph[PIX(-1, dy, dz)];
                    Ph[PIX(-1, dy, dz)] = 0;
                }
                if (np & XHI) {
                    Ph[PIX(0, dy, dz)] += Ph[PIX(1, dy, dz)];
                    Ph[PIX(1, dy, dz)] = 0;
                }
            }
    }
    if (np & (YLO | YHI)) {
#pragma unroll
        for (int dx = -1; dx <= 1; dx++)
#pragma unroll
            for (int dz = -1; dz <= 1; dz++) {
                if (np & YLO) {
                    Ph[PIX(dx, 0, dz)] += Ph[PIX(dx, -1, dz)];
                    Ph[PIX(dx, -1, dz)] = 0;
                }
                if (np & YHI) {
                    Ph[PIX(dx, 0, dz)] += Ph[PIX(dx, 1, dz)];
                    Ph[PIX(dx, 1, dz)] = 0;
                }
            }
    }
    if (np & (ZLO | ZHI)) {
#pragma unroll
        for (int dx = -1; dx <= 1; dx++)
#pragma unroll
            for (int dy = -1; dy <= 1; dy++) {
                if (np & ZLO) {
                    Ph[PIX(dx, dy, 0)] += Ph[PIX(dx, dy, -1)];
                    Ph[PIX(dx, dy, -1)] = 0;
                }
                if (np & ZHI) {
                    Ph[PIX(dx, dy, 0)] += Ph[PIX(dx, dy, 1)];
                    Ph[PIX(dx, dy, 1)] = 0;
                }
            }
    }
    // the columns with dz != 0 parts: (0,0), (-1,0), (1,0), (0,-1), (0,1)
    const uint32_t tid = threadIdx.x;
    constexpr int CX[5] = {0, -1, 1, 0, 0}, CY[5] = {0, 0, 0, -1, 1};
#pragma unroll
    for (int k = 0; k < 5; k++) {
        zq[2 * k][tid] = Ph[PIX(CX[k], CY[k], 1)];       // to z + 1
        zq[2 * k + 1][tid] = Ph[PIX(CX[k], CY[k], -1)];  // to z - 1
    }
    __syncthreads();
    if (!act) return;
    const uint32_t zm1 = (uint32_t)a.nz - 1u;
    const uint32_t rb = tid - (tid & zm1);
    const uint32_t below = rb + ((uint32_t)(c.z - 1) & zm1), above = rb + ((uint32_t)(c.z + 1) & zm1);
    const size_t N = a.N;
#pragma unroll
    for (int dx = -1; dx <= 1; dx++)
#pragma unroll
        for (int dy = -1; dy <= 1; dy++) {
            R v = Ph[PIX(dx, dy, 0)];
#pragma unroll
            for (int k = 0; k < 5; k++)
                if (CX[k] == dx && CY[k] == dy) v += zq[2 * k][below] + zq[2 * k + 1][above];
            stp(slot_at(Q + (size_t)((dx + 1) * 3 + (dy + 1)) * N, nbr(c, dx, dy, 0)), v);
        }
}

// ---------------------------------------------------------------- kernels
// Minimum waves per SIMD of the FP32 reverse kernel (the register bound follows from it).
#ifndef MKG_ADJ_WAVES_F32
#define MKG_ADJ_WAVES_F32 4
#endif
// FLOW: the coupled model.  The scalars are advected with u~ = (1 - n_s) u, u the flow
// velocity at level n and n_s the solid fraction of partial bounce-back:
//   h_eq,k = w_k T (1 + c_k.u~ / cs_h^2),  g_eq,i = w_i (phi + (phi + 1) c_i.u~ / cs_g^2),
// so their reverse adds w_k (c_k.u~) / cs^2 terms to dL/dT and dL/dphi and emits dL/du~
// (j.ut) for the flow kernel; the buoyancy part of dL/dT arrives as a cell-wide thermal
// adjoint (j.hcw_in) with the gathered populations.
// GQ: the block holds whole z rows.  The stencil contributions of a cell are summed per
// (dx, dy) column of the stencil into nine row sums, the z shifts taken through LDS, and
// written to nine N-arrays at the target rows (see group_stencil_adjoint); otherwise the 19
// contributions are written per direction for k_adj_stencil.
template <class R, bool FLOW, bool GQ>
__global__ void __launch_bounds__(256, sizeof(R) == 4 ? MKG_ADJ_WAVES_F32 : 1)
    k_adj_collide(const StepArgs a, const AdjArgs j) {
    Cell c;
    const bool valid = GQ ? locate_clamped(a, c) : locate(a, c);
    const uint32_t N = a.N;
    const uint32_t id = c.id;
    __shared__ double red[4];
    __shared__ R zq[GQ ? 10 : 1][256];
    double eps_loc = 0;
    R Ph[27];
#pragma unroll
    for (int k = 0; k < 27; k++) Ph[k] = 0;
    if (valid && a.flag[id] != kIsWalls) {
        R gs[19], hs[7];
        const R* hcw = FLOW ? (const R*)j.hcw_in : nullptr;
        if ((c.bits & (uint32_t)(a.g.npmask | a.h.npmask)) == 0u) {
            adj_gather<19, R, false>(gs, (const R*)j.gin, (const R*)j.nin, a.g, c, a, kKindScalar, N);
            adj_gather<7, R, false>(hs, (const R*)j.hin, hcw, a.h, c, a, kKindScalar, N);
        } else {
            adj_gather<19, R>(gs, (const R*)j.gin, (const R*)j.nin, a.g, c, a, kKindScalar, N);
            adj_gather<7, R>(hs, (const R*)j.hin, hcw, a.h, c, a, kKindScalar, N);
        }
        // forward values at level n
        R h[7];
        for (int k = 0; k < 7; k++) h[k] = *slot_at((const R*)j.htape + (size_t)k * a.S, c.id);
        R sh = 0;
#pragma unroll
        for (int k = 0; k < 7; k++) sh += h[k];
        const R T = (R)a.h.ref + sh;
        const R* __restrict__ ph = (const R*)j.phin;
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
        R P[27];
#pragma unroll
        for (int k = 0; k < 27; k++) P[k] = 0;
#pragma unroll
        for (int i = 0; i < 19; i++)
            P[PIX(cx(i), cy(i), cz(i))] = *slot_at(ph, px[cx(i) + 1] + py[cy(i) + 1] + pz[cz(i) + 1]);
        const R phi = P[PIX(0, 0, 0)];
        R Mrot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        bool rot = false;
        if (a.ngrains > 0) {
            const int gid = a.grain[id];
            if (gid > 0) {
#pragma unroll
                for (int k = 0; k < 9; k++) Mrot[k] = (R)a.rot[(gid - 1) * 9 + k];
                rot = true;
            }
        }
        // thermal rate and its derivatives
        const R fsr = (R)0.5 * ((R)1 + phi);
        const R fs = fsr < (R)0 ? (R)0 : (fsr > (R)1 ? (R)1 : fsr);
        const R dalpha = (fsr > (R)0 && fsr < (R)1) ? (R)0.5 * ((R)a.alpha_s - (R)a.alpha_l) : (R)0;
        const R alpha0 = (R)a.alpha_l * ((R)1 - fs) + (R)a.alpha_s * fs;
        const R kap = a.has_kappa ? ((const R*)a.kappa)[id] : (R)1;
        const R om = (R)1 / ((R)0.5 + kap * alpha0 / (R)0.25);
        const R dom_dphi = -om * om * kap * dalpha / (R)0.25;
        const R dom_dkap = -om * om * alpha0 / (R)0.25;
        // advection velocity of the scalars at level n (coupled model): u from the flow tape
        // with the body force and buoyancy, times the liquid fraction 1 - n_s
        R v0 = 0, v1 = 0, v2 = 0;
        if constexpr (FLOW) {
            R rho = 0, j0 = 0, j1 = 0, j2 = 0;
#pragma unroll
            for (int i = 0; i < 19; i++) {
                const R fi = (R)j.ftape[(size_t)i * N + id];
                rho += fi;
                if (cx(i)) j0 += (R)cx(i) * fi;
                if (cy(i)) j1 += (R)cy(i) * fi;
                if (cz(i)) j2 += (R)cz(i) * fi;
            }
            const R dT = T - (R)a.T_ref;
            const R F0 = (R)a.force[0] + (R)(a.rho0 * a.buoy[0]) * dT;
            const R F1 = (R)a.force[1] + (R)(a.rho0 * a.buoy[1]) * dT;
            const R F2 = (R)a.force[2] + (R)(a.rho0 * a.buoy[2]) * dT;
            // the flag of level n, from phi^n of the tape as the forward's moment pass set it
            // (the flag array itself may hold another level while a segment is recomputed)
            const int8_t fl = classify(a.flag[id], (double)phi, a.phi_s, a.pf_model);
            R ns = 0;
            if (fl == kIsSolid) ns = (R)1;
            else if (fl == kIsInter) ns = fs;
            const R q = ((R)1 - ns) / rho;
            v0 = (j0 + (R)0.5 * F0) * q;
            v1 = (j1 + (R)0.5 * F1) * q;
            v2 = (j2 + (R)0.5 * F2) * q;
        }
        // adjoint sums; with advection, B and A carry the (c.u~)/cs^2 part of the equilibria
        R A = 0, B = 0, Om = 0, Au = 0;
        R ut0 = 0, ut1 = 0, ut2 = 0;  // dL/du~ of the two scalar equilibria
#pragma unroll
        for (int i = 0; i < 19; i++) {
            const R wl = (R)wgt<19>(i) * gs[i];
            A += wl;
            if constexpr (FLOW) {
                const R cv = cx(i) * v0 + cy(i) * v1 + cz(i) * v2;
                Au += (R)3 * cv * wl;
                if (cx(i)) ut0 += (R)(3 * cx(i)) * wl;
                if (cy(i)) ut1 += (R)(3 * cy(i)) * wl;
                if (cz(i)) ut2 += (R)(3 * cz(i)) * wl;
            }
        }
        R uh0 = 0, uh1 = 0, uh2 = 0;
#pragma unroll
        for (int k = 0; k < 7; k++) {
            const R wl = (R)wgt<7>(k) * hs[k];
            R heq = (R)wgt<7>(k) * sh;
            if constexpr (FLOW) {
                const R cv = cx(k) * v0 + cy(k) * v1 + cz(k) * v2;
                heq += (R)wgt<7>(k) * T * (R)4 * cv;
                B += wl * ((R)1 + (R)4 * cv);
                if (cx(k)) uh0 += (R)(4 * cx(k)) * wl;
                if (cy(k)) uh1 += (R)(4 * cy(k)) * wl;
                if (cz(k)) uh2 += (R)(4 * cz(k)) * wl;
            } else {
                B += wl;
            }
            Om -= (h[k] - heq) * hs[k];
        }
        R eh = 0;
        const R upf = a.pf_coupling == 1 ? T : (R)0;
        const R uh = phase_source_adj<R>(P, upf, (R)a.eps4, (R)a.lambda, (R)a.pf_dt,
                                         (R)a.pf_wdx2, Mrot, rot, A, Ph, eh);
        const R Th = om * B + (a.pf_coupling == 1 ? uh : (R)0);
        R* __restrict__ ho = (R*)j.hout;
#pragma unroll
        for (int k = 0; k < 7; k++) stp(slot_at(ho + (size_t)k * N, id), ((R)1 - om) * hs[k] + Th);
        const R L = (R)a.latent_T;
        const R pin = ((const R*)j.pin)[id];
        // the latent source w_k L dphi has no advective part: its adjoint is L sum_k w_k hs_k
        R Bl = B;
        if constexpr (FLOW) {
            Bl = 0;
#pragma unroll
            for (int k = 0; k < 7; k++) Bl += (R)wgt<7>(k) * hs[k];
        }
        const R phl = (R)a.omega_g * (A + Au) + L * Bl + pin + Om * dom_dphi;
        if constexpr (FLOW) {
            // dL/du~ = omega_g sum_i w_i gs_i (phi + 1) 3 c_i + omega_h sum_k w_k hs_k T 4 c_k
            R* __restrict__ ut = (R*)j.ut;
            const R pg = (R)a.omega_g * (phi + (R)1), ph_ = om * T;
            ut[id] = pg * ut0 + ph_ * uh0;
            ut[(size_t)N + id] = pg * ut1 + ph_ * uh1;
            ut[2 * (size_t)N + id] = pg * ut2 + ph_ * uh2;
        }
        R* __restrict__ go = (R*)j.gout;
        const R og = (R)1 - (R)a.omega_g;
#pragma unroll
        for (int i = 0; i < 19; i++) stp(slot_at(go + (size_t)i * N, id), og * gs[i] + phl);
        ((R*)j.pout)[id] = -L * Bl;
        if (a.has_kappa) ((R*)j.kap_grad)[id] += Om * dom_dkap;
        if constexpr (!GQ) {
            R* __restrict__ Pb = (R*)j.P;
#pragma unroll
            for (int i = 0; i < 19; i++)
                stp(slot_at(Pb + (size_t)i * N, id), Ph[PIX(cx(i), cy(i), cz(i))]);
        }
        eps_loc = (double)eh;
    }
    if constexpr (GQ) group_stencil_adjoint<R>(a, c, valid, Ph, zq, (R*)j.P);
    // FP64 block sum of dL/deps4 (wavefront butterfly + LDS), accumulated per block
    double v = eps_loc;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t b = blockIdx.x;
        j.eps_part[b] += red[0] + red[1] + red[2] + red[3];
    }
}

// Transpose of the clamped / periodic 19-point stencil: phi(x) was read by cell y through
// direction k when wrapclamp(y + c_k) == x.
template <class R>
__global__ void __launch_bounds__(256) k_adj_stencil(const StepArgs a, const R* __restrict__ P,
                                                     R* __restrict__ nout) {
    Cell c;
    if (!locate(a, c)) return;
    const uint32_t N = a.N;
    const uint32_t np = (uint32_t)a.g.npmask;
    R s = 0;
    if ((c.bits & np) == 0u) {
        // away from non-periodic faces every direction has its one regular source
#pragma unroll
        for (int k = 0; k < 19; k++)
            s += ldp(slot_at(P + (size_t)k * N, nbr(c, -cx(k), -cy(k), -cz(k))));
        nout[c.id] = s;
        return;
    }
#pragma unroll
    for (int k = 0; k < 19; k++) {
        // per axis: the regular source (x - d, invalid across a non-periodic face) and, on a
        // non-periodic face, the cell itself read through the clamp
        uint32_t ox[2], oy[2], oz[2];
        bool vx[2], vy[2], vz[2];
        auto axis = [&](int d, const uint32_t* off, uint32_t lo, uint32_t hi, uint32_t* o,
                        bool* v) {
            if (d == 0) {
                o[0] = off[1];
                v[0] = true;
                o[1] = off[1];
                v[1] = false;
            } else {
                const uint32_t back = d > 0 ? lo : hi;
                const uint32_t fwd = d > 0 ? hi : lo;
                o[0] = off[1 - d];
                v[0] = !((c.bits & back) && (np & back));
                o[1] = off[1];
                v[1] = (c.bits & fwd) && (np & fwd);
            }
        };
        axis(cx(k), c.ox, XLO, XHI, ox, vx);
        axis(cy(k), c.oy, YLO, YHI, oy, vy);
        axis(cz(k), c.oz, ZLO, ZHI, oz, vz);
#pragma unroll
        for (int i0 = 0; i0 < 2; i0++)
#pragma unroll
            for (int i1 = 0; i1 < 2; i1++)
#pragma unroll
                for (int i2 = 0; i2 < 2; i2++)
                    if (vx[i0] && vy[i1] && vz[i2])
                        s += ldp(slot_at(P + (size_t)k * N, ox[i0] + oy[i1] + oz[i2]));
    }
    nout[c.id] = s;
}

// The phi adjoint of every cell from the nine column sums of group_stencil_adjoint.
template <class R>
__global__ void __launch_bounds__(256) k_adj_qsum(const StepArgs a, const R* __restrict__ Q,
                                                  R* __restrict__ nout) {
    const uint32_t t = thread_cell(a);
    if (t >= a.N) return;
    const size_t N = a.N;
    R s = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) s += ldp(slot_at(Q + (size_t)k * N, t));
    nout[t] = s;
}

// Loss L = 1/2 sum (phi - target)^2 (FP64 block partials) and the seed of the reverse sweep:
// the cell-wide phase adjoint phi - target.
template <class R>
__global__ void __launch_bounds__(256) k_adj_loss(const StepArgs a, const double* target,
                                                  R* __restrict__ nseed, double* part) {
    Cell c;
    __shared__ double red[4];
    double v = 0;
    if (locate(a, c) && a.flag[c.id] != kIsWalls) {
        R g[19];
        read_pops_rt<19, R>(g, a.g, c, a, kKindScalar);
        R s = 0;
#pragma unroll
        for (int i = 0; i < 19; i++) s += g[i];
        const double phi = a.g.ref + (double)s;
        const double d = phi - target[c.id];
        nseed[c.id] = (R)d;
        v = 0.5 * d * d;
    } else if (locate(a, c)) {
        nseed[c.id] = 0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void k_sum_partials(const double* part, uint32_t n, double* out) {
    __shared__ double red[4];
    double v = 0;
    for (uint32_t i = threadIdx.x; i < n; i += 256) v += part[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) *out = red[0] + red[1] + red[2] + red[3];
}

// Flow reverse of the coupled step (after k_adj_collide<R, true> and k_adj_stencil), per
// non-wall cell y.  Forward: rho, u = (j + F/2)/rho with F = force + rho0 buoy (T - T_ref);
// post = the generalised moment collision with Guo forcing; out_i = (1 - n_s) post_i
// + n_s f_inv(i) (partial bounce-back); out streams with the face rules of the flow lattice;
// the scalars are advected with u~ = (1 - n_s) u.  Reverse:
//   l = the streamed flow adjoint of level n+1 (the gather of the flow-only reverse),
//   dL/df = collide_reverse((1 - n_s) l; velocity adjoint (1 - n_s) dL/du~) + n_s l_inv,
//   dL/dn_s = sum_i l_i (f_inv(i) - post_i) - u . dL/du~ (+ the moving-wall terms of bounce
//             faces, which scale with 1 - n_s),
//   dL/dT (buoyancy) = rho0 buoy . dL/dF   -> the cell-wide thermal adjoint of level n,
//   dL/dphi += dL/dn_s / 2 in interface cells (n_s = (1 + phi)/2 there) -> the cell-wide phase
//   adjoint of level n (added to what k_adj_stencil wrote).
template <class R>
__global__ void __launch_bounds__(256, sizeof(R) == 4 ? 2 : 1)
    k_cadj_flow(const StepArgs a, const AdjArgs j) {
    constexpr int Q = 19;
    const uint32_t N = a.N;
    Cell c;
    if (!locate(a, c)) return;
    const uint32_t id = c.id;
    const R* __restrict__ lam_in = (const R*)j.fin;
    R* __restrict__ lam_out = (R*)j.fout;
    if (a.flag[id] == kIsWalls) {
        for (int i = 0; i < Q; i++) lam_out[(size_t)i * N + id] = 0;
        ((R*)j.hcw_out)[id] = 0;
        return;
    }
    // l_i(y) = lambda_i(y + c_i), reflected at bounce faces and walls, plus the outflow copy.
    // A population returned by a bounce face carries the moving-wall term
    // 6 w rho0 c_inv(i).u_w (1 - n_s): its n_s adjoint collects in dns_wall.
    const uint32_t np = (uint32_t)a.f.npmask & c.bits;
    R l[Q];
    R dns_wall = 0;
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const int X = cx(i), Y = cy(i), Z = cz(i);
        const uint32_t hp = np & faces_of(X, Y, Z);
        R v = 0;
        if (hp == 0u) {
            const uint32_t t = nbr(c, X, Y, Z);
            v = a.flag[t] == kIsWalls ? *slot_at(lam_in + (size_t)inv(i) * N, id)
                                      : *slot_at(lam_in + (size_t)i * N, t);
        } else if (pick6(__builtin_ctz(hp), a.f.bc) == kBcBounce) {
            v = *slot_at(lam_in + (size_t)inv(i) * N, id);
            const double cu = wall_cu(a, __builtin_ctz(hp), c, X, Y, Z);
            dns_wall += v * (R)(6.0 * wgt<Q>(i) * a.rho0 * cu);
        }
        const uint32_t hb = np & faces_of(-X, -Y, -Z);
        if (hb != 0u && pick6(__builtin_ctz(hb), a.f.bc) == kBcOutflow)
            v += *slot_at(lam_in + (size_t)i * N, id);
        l[i] = v;
    }
    // forward values at level n
    R f[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) f[i] = (R)j.ftape[(size_t)i * N + id];
    R h[7];
    // the tape holds the pre-collision thermal populations slot by slot (stride S)
    for (int k = 0; k < 7; k++) h[k] = *slot_at((const R*)j.htape + (size_t)k * a.S, c.id);
    R sh = 0;
#pragma unroll
    for (int k = 0; k < 7; k++) sh += h[k];
    const R T = (R)a.h.ref + sh;
    const R phi = *slot_at((const R*)j.phin, id);
    // the flag of level n from phi^n, as in k_adj_collide
    const int8_t fl = classify(a.flag[id], (double)phi, a.phi_s, a.pf_model);
    R ns = 0, dns_dphi = 0;
    if (fl == kIsSolid) {
        ns = (R)1;
    } else if (fl == kIsInter) {
        const R q = (R)0.5 * ((R)1 + phi);
        ns = q < (R)0 ? (R)0 : (q > (R)1 ? (R)1 : q);
        dns_dphi = (q > (R)0 && q < (R)1) ? (R)0.5 : (R)0;
    }
    const R dT = T - (R)a.T_ref;
    const R F[3] = {(R)a.force[0] + (R)(a.rho0 * a.buoy[0]) * dT,
                    (R)a.force[1] + (R)(a.rho0 * a.buoy[1]) * dT,
                    (R)a.force[2] + (R)(a.rho0 * a.buoy[2]) * dT};
    R rho = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int i = 0; i < Q; i++) {
        rho += f[i];
        if (cx(i)) jx += (R)cx(i) * f[i];
        if (cy(i)) jy += (R)cy(i) * f[i];
        if (cz(i)) jz += (R)cz(i) * f[i];
    }
    const R ux = (jx + (R)0.5 * F[0]) / rho, uy = (jy + (R)0.5 * F[1]) / rho,
            uz = (jz + (R)0.5 * F[2]) / rho;
    // post-collision populations without the partial bounce-back (full values: reference 0)
    R post[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) post[i] = f[i];
    flow_collide<Q, R>(post, rho, rho, ux, uy, uz, F[0], F[1], F[2], (R)a.s_nu, (R)a.s_bulk,
                       (R)a.s_odd, (R)a.s_even, (R)0);
    const R keep = (R)1 - ns;
    R dns = dns_wall, lp[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) {
        dns += l[i] * (f[inv(i)] - post[i]);
        lp[i] = keep * l[i];
    }
    const R* __restrict__ ut = (const R*)j.ut;
    const R w0 = *slot_at(ut, id), w1 = *slot_at(ut + (size_t)N, id), w2 = *slot_at(ut + 2 * (size_t)N, id);
    dns -= ux * w0 + uy * w1 + uz * w2;
    const R ub_ext[3] = {keep * w0, keep * w1, keep * w2};
    R dF[3], drate[4];
    collide_reverse<Q, R, false>(a, f, F, lp, ub_ext, lam_out, N, id, dF, drate);
    // partial bounce-back: out_i takes n_s f_inv(i)
#pragma unroll
    for (int i = 0; i < Q; i++) {
        R* o = &lam_out[(size_t)i * N + id];
        *o = *o + ns * l[inv(i)];
    }
    ((R*)j.hcw_out)[id] = (R)a.rho0 * ((R)a.buoy[0] * dF[0] + (R)a.buoy[1] * dF[1] +
                                       (R)a.buoy[2] * dF[2]);
    R* nacc = (R*)j.nacc;
    nacc[id] = nacc[id] + dns_dphi * dns;
}

// ---------------------------------------------------------------- launchers
int launch_adj_collide(int fp32, int grouped, const StepArgs& a, const AdjArgs& j,
                       hipStream_t s) {
    if (grouped) {
        if (fp32) k_adj_collide<float, false, true><<<a.nblk, 256, 0, s>>>(a, j);
        else k_adj_collide<double, false, true><<<a.nblk, 256, 0, s>>>(a, j);
    } else {
        if (fp32) k_adj_collide<float, false, false><<<a.nblk, 256, 0, s>>>(a, j);
        else k_adj_collide<double, false, false><<<a.nblk, 256, 0, s>>>(a, j);
    }
    return 0;
}
int launch_adj_collide_flow(int fp32, int grouped, const StepArgs& a, const AdjArgs& j,
                            hipStream_t s) {
    if (grouped) {
        if (fp32) k_adj_collide<float, true, true><<<a.nblk, 256, 0, s>>>(a, j);
        else k_adj_collide<double, true, true><<<a.nblk, 256, 0, s>>>(a, j);
    } else {
        if (fp32) k_adj_collide<float, true, false><<<a.nblk, 256, 0, s>>>(a, j);
        else k_adj_collide<double, true, false><<<a.nblk, 256, 0, s>>>(a, j);
    }
    return 0;
}
int launch_adj_qsum(int fp32, const StepArgs& a, const void* Q, void* nout, hipStream_t s) {
    if (fp32) k_adj_qsum<float><<<a.nblk, 256, 0, s>>>(a, (const float*)Q, (float*)nout);
    else k_adj_qsum<double><<<a.nblk, 256, 0, s>>>(a, (const double*)Q, (double*)nout);
    return 0;
}
int launch_cadj_flow(int fp32, const StepArgs& a, const AdjArgs& j, hipStream_t s) {
    if (fp32) k_cadj_flow<float><<<a.nblk, 256, 0, s>>>(a, j);
    else k_cadj_flow<double><<<a.nblk, 256, 0, s>>>(a, j);
    return 0;
}
int launch_adj_stencil(int fp32, const StepArgs& a, const void* P, void* nout, hipStream_t s) {
    if (fp32) k_adj_stencil<float><<<a.nblk, 256, 0, s>>>(a, (const float*)P, (float*)nout);
    else k_adj_stencil<double><<<a.nblk, 256, 0, s>>>(a, (const double*)P, (double*)nout);
    return 0;
}
int launch_adj_loss(int fp32, const StepArgs& a, const double* target, void* nseed, double* part,
                    double* out, hipStream_t s) {
    if (fp32) k_adj_loss<float><<<a.nblk, 256, 0, s>>>(a, target, (float*)nseed, part);
    else k_adj_loss<double><<<a.nblk, 256, 0, s>>>(a, target, (double*)nseed, part);
    k_sum_partials<<<1, 256, 0, s>>>(part, a.nblk, out);
    return 0;
}
int launch_sum_partials(const double* part, uint32_t n, double* out, hipStream_t s) {
    k_sum_partials<<<1, 256, 0, s>>>(part, n, out);
    return 0;
}

}  // namespace mkg
