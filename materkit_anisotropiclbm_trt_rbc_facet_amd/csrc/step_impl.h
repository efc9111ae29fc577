// Fused time-step kernels.  Included by step_f32.hip / step_f64.hip, which instantiate them
// for one storage precision each.
//
// One time step of the coupled solver is two launches:
//   k_phase_moments  phi^n = sum_i g_i for every cell (needed by the neighbours' gradient
//                    stencils before any phase collision), flags re-classified from phi^n
//                    (mkFluidsUser::relaxation first sweep, MaterKit/src/mkFluidsUser.cpp:
//                    147-172, and mkPhaseField::reset_flags, MaterKit/inc/mkPhaseField.h:37)
//   k_step           per cell: thermal / solute moments, flow collide-stream with Guo forcing,
//                    Boussinesq buoyancy and partial bounce-back in solid, thermal and
//                    solute collide-stream with the latent / rejection source, phase-field
//                    collide-stream with the anisotropic source -- every lattice in one pass
//                    (the per-field propagate()/relaxation() pairs of MaterKit.cpp:500-518)
// Without a phase lattice only k_step runs.
#pragma once
#include "device.h"

namespace mkg {

template <class R>
__device__ inline R clamp01(R v) {
    return v < (R)0 ? (R)0 : (v > (R)1 ? (R)1 : v);
}

// 19-point phase stencil around the cell (periodic faces wrap, other faces replicate the
// edge value), index (dx+1)*9 + (dy+1)*3 + (dz+1).
template <class R>
__device__ inline void load_phase_stencil(const StepArgs& a, const Cell& c, R (&P)[27]) {
    const R* __restrict__ ph = (const R*)a.phi_cur;
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
            ph[px[cx(i) + 1] + py[cy(i) + 1] + pz[cz(i) + 1]];
}

// Isotropic LB-weighted gradient of the stencil (D3Q19 weights over cs^2 = 1/3).
template <class R>
__device__ inline void stencil_gradient(const R (&P)[27], R (&gr)[3]) {
    gr[0] = gr[1] = gr[2] = 0;
#pragma unroll
    for (int i = 1; i < 19; i++) {
        const R v = (R)(3.0 * wgt<19>(i)) * P[(cx(i) + 1) * 9 + (cy(i) + 1) * 3 + (cz(i) + 1)];
        if (cx(i)) gr[0] += (R)cx(i) * v;
        if (cy(i)) gr[1] += (R)cy(i) * v;
        if (cz(i)) gr[2] += (R)cz(i) * v;
    }
}

template <class R, int PQ, bool ODD>
__global__ void __launch_bounds__(256) k_phase_moments(const StepArgs a, R* __restrict__ phi_out) {
    Cell c;
    if (!locate(a, c)) return;
    const int8_t fl = a.flag[c.id];
    if (fl == kIsWalls) {
        // wall cells keep their phase value: carry it into the buffer being written
        phi_out[c.id] = ((const R*)a.phi_prev)[c.id];
        return;
    }
    R g[PQ];
    read_pops<PQ, ODD, R>(g, a.g, a.S, c, a, kKindScalar);
    R s = 0;
#pragma unroll
    for (int i = 0; i < PQ; i++) s += g[i];
    const R phi = (R)a.g.ref + s;
    phi_out[c.id] = phi;
    const int8_t nf = classify(fl, (double)phi, a.phi_s);
    if (nf != fl) a.flag[c.id] = nf;
}

// Minimum waves per SIMD requested from the register allocator for the fused kernel.  FP32:
// four (at most 128 VGPRs).  Left alone, the allocator gives the z-staged O step 136 VGPRs and
// three waves; with four it spills 12 bytes per lane and the step runs 7 % faster at 256^3.
// FP64 kernels need about 230 VGPRs (two waves); a bound there only adds spills.
#ifndef MKG_STEP_WAVES_F32
#define MKG_STEP_WAVES_F32 4
#endif

// LDS slots of the z stage shared by the flow and phase lattices (one after the other).
template <int FQ, int PQ>
constexpr int zstage_slots() {
    int n = 1;
    if (FQ == 19 && nzdirs<19>() > n) n = nzdirs<19>();
    if (FQ == 27 && nzdirs<27>() > n) n = nzdirs<27>();
    if (PQ == 19 && nzdirs<19>() > n) n = nzdirs<19>();
    if (PQ == 7 && nzdirs<7>() > n) n = nzdirs<7>();
    return n;
}

// Population access of one lattice in k_step: E step, O step, or O step with the z stage.
template <int Q, bool ODD, bool Z, class R, int ZN>
__device__ inline void pull(R* p, const LatArgs& L, uint32_t S, const Cell& c, const StepArgs& a,
                            int kind, R (*zv)[256], uint32_t tid) {
    if constexpr (Z)
        read_pops_z<Q, R, ZN>(p, L, S, c, a, zv, tid);
    else
        read_pops<Q, ODD, R>(p, L, S, c, a, kind);
}
template <int Q, bool ODD, bool Z, class R, int ZN>
__device__ inline void push(const R* p, const LatArgs& L, uint32_t S, const Cell& c,
                            const StepArgs& a, int kind, R (*zv)[256], uint32_t* zm,
                            uint32_t tid, bool act, const FaceState<R>& fs) {
    if constexpr (Z)
        write_pops_z<Q, R, ZN>(p, L, S, c, a, kind, zv, zm, tid, act, fs);
    else
        write_pops<Q, ODD, R>(p, L, S, c, a, kind, fs);
}

// ZS: O step with the z stage (only instantiated with ODD; the block must hold whole z rows).
// With it every thread stays to the end (barriers), and act marks the threads that update.
// Each pull and push has its own LDS region, except that the phase lattice reuses the flow's
// regions: every thread has passed the flow push's barrier (after its last read of the flow
// pull region) and the phase pull's barrier (after its last read of the flow push region)
// before it writes them again.
template <class R, int FQ, int PQ, bool TH, bool SO, bool ODD, bool ZS = false>
__global__ void __launch_bounds__(256, sizeof(R) == 4 ? MKG_STEP_WAVES_F32 : 1)
    k_step(const StepArgs a) {
    constexpr bool Z = ODD && ZS;
    constexpr int ZN = Z ? zstage_slots<FQ, PQ>() : 1;
    constexpr int ZH = (Z && TH) ? nzdirs<7>() : 1;
    constexpr int ZC = (Z && SO) ? nzdirs<7>() : 1;
    __shared__ R zrf[ZN][256], zwf[ZN][256];  // flow, then phase: pull / push
    __shared__ R zrh[ZH][256], zwh[ZH][256];  // thermal
    __shared__ R zrc[ZC][256], zwc[ZC][256];  // solute
    __shared__ uint32_t zmf[Z ? 256 : 1], zmh[(Z && TH) ? 256 : 1], zmc[(Z && SO) ? 256 : 1];
    const uint32_t tid = threadIdx.x;
    Cell c;
    bool act = true;
    if constexpr (Z) {
        act = locate_clamped(a, c);
    } else {
        if (!locate(a, c)) return;
    }
    const int8_t fl = a.flag[c.id];
    if constexpr (Z) {
        act = act && fl != kIsWalls;
    } else {
        if (fl == kIsWalls) return;
    }
    const uint32_t S = a.S;

    R phi = (R)-1, dphi = (R)0, ns = (R)0;
    if constexpr (PQ > 0) {
        phi = ((const R*)a.phi_cur)[c.id];
        dphi = phi - ((const R*)a.phi_prev)[c.id];
        if (fl == kIsSolid)
            ns = (R)1;
        else if (fl == kIsInter)
            ns = clamp01((R)0.5 * ((R)1 + phi));
    }

    // ---- scalar moments at time level n
    R h[TH ? 7 : 1];
    R sh = 0, T = 0;
    if constexpr (TH) {
        if (a.h.adv)
            pull<7, ODD, Z, R, ZH>(h, a.h, S, c, a, kKindScalar, zrh, tid);
        else
            read_pops_rt<7, R>(h, a.h, S, c, a, kKindScalar);
#pragma unroll
        for (int i = 0; i < 7; i++) sh += h[i];
        T = (R)a.h.ref + sh;
    }
    R cc[SO ? 7 : 1];
    R sc = 0, C = 0, U = 0;
    if constexpr (SO) {
        if (a.c.adv)
            pull<7, ODD, Z, R, ZC>(cc, a.c, S, c, a, kKindScalar, zrc, tid);
        else
            read_pops_rt<7, R>(cc, a.c, S, c, a, kKindScalar);
#pragma unroll
        for (int i = 0; i < 7; i++) sc += cc[i];
        C = (R)a.c.ref + sc;
        const R A = (R)0.5 * (((R)1 + (R)a.k_part) - ((R)1 - (R)a.k_part) * phi);
        U = C / A;
    }

    // ---- flow
    R ux = 0, uy = 0, uz = 0;
    if constexpr (FQ > 0) {
        R f[FQ];
        if (a.f.adv)
            pull<FQ, ODD, Z, R, ZN>(f, a.f, S, c, a, kKindFlow, zrf, tid);
        else
            read_pops_rt<FQ, R>(f, a.f, S, c, a, kKindFlow);
        R dr = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
        for (int i = 0; i < FQ; i++) {
            dr += f[i];
            if (cx(i)) jx += (R)cx(i) * f[i];
            if (cy(i)) jy += (R)cy(i) * f[i];
            if (cz(i)) jz += (R)cz(i) * f[i];
        }
        const R rho = (R)a.f.ref + dr;
        R Fx = (R)a.force[0], Fy = (R)a.force[1], Fz = (R)a.force[2];
        if constexpr (TH) {
            const R dT = T - (R)a.T_ref;
            Fx += (R)(a.rho0 * a.buoy[0]) * dT;
            Fy += (R)(a.rho0 * a.buoy[1]) * dT;
            Fz += (R)(a.rho0 * a.buoy[2]) * dT;
        }
        // the density in the u terms: rho, or rho0 for the incompressible equilibrium
        const R den = a.incomp ? (R)a.rho0 : rho;
        ux = (jx + (R)0.5 * Fx) / den;
        uy = (jy + (R)0.5 * Fy) / den;
        uz = (jz + (R)0.5 * Fz) / den;
        if (a.f.adv) {
            flow_collide<FQ, R>(f, dr, den, ux, uy, uz, Fx, Fy, Fz, (R)a.s_nu, (R)a.s_bulk,
                                (R)a.s_odd, (R)a.s_even, ns);
            FaceState<R> fs;
            fs.wscale = (R)1 - ns;
            fs.ux = ux;
            fs.uy = uy;
            fs.uz = uz;
            push<FQ, ODD, Z, R, ZN>(f, a.f, S, c, a, kKindFlow, zwf, zmf, tid, act, fs);
        }
        const R om = (R)1 - ns;
        ux *= om;
        uy *= om;
        uz *= om;
    }

    // ---- thermal collide-stream (mkThermalField::propagate + relaxation)
    if constexpr (TH) {
        if (a.h.adv) {
            const R fs = PQ > 0 ? clamp01((R)0.5 * ((R)1 + phi)) : (R)0;
            R alpha = (R)a.alpha_l * ((R)1 - fs) + (R)a.alpha_s * fs;
            if (a.has_kappa) alpha *= ((const R*)a.kappa)[c.id];
            const R om = (R)1 / ((R)0.5 + alpha / (R)0.25);
            const R src = (R)a.latent_T * dphi;
#pragma unroll
            for (int i = 0; i < 7; i++) {
                const R w = (R)wgt<7>(i);
                const R cu = cx(i) * ux + cy(i) * uy + cz(i) * uz;
                const R heq = w * (sh + T * cu / (R)0.25);
                h[i] = h[i] - om * (h[i] - heq) + w * src;
            }
            push<7, ODD, Z, R, ZH>(h, a.h, S, c, a, kKindScalar, zwh, zmh, tid, act,
                                   FaceState<R>());
        }
    }

    // ---- solute collide-stream (mkSolutalField::propagate + relaxation): the conserved
    //      c = A(phi) U obeys dc/dt + div(c u) = div(D q grad U), written as an
    //      advection-diffusion of c with D_c = D q / A and the drift V = D q grad A / A^2
    R P[27];
    if constexpr (SO && PQ > 0) load_phase_stencil<R>(a, c, P);
    if constexpr (SO) {
        if (a.c.adv) {
            R gphi[3] = {0, 0, 0};
            if constexpr (PQ > 0) stencil_gradient<R>(P, gphi);
            R q = (R)0.5 * ((R)1 - phi);
            if (a.D_l > 0) q += (R)(a.D_s / a.D_l) * (R)0.5 * ((R)1 + phi);
            if (q < (R)a.q_min) q = (R)a.q_min;
            const R A = (R)0.5 * (((R)1 + (R)a.k_part) - ((R)1 - (R)a.k_part) * phi);
            const R Dc = (R)a.D_l * q / A;
            const R om = (R)1 / ((R)0.5 + Dc / (R)0.25);
            const R vfac = (R)(-0.5 * (1.0 - a.k_part)) * (R)a.D_l * q / (A * A);
            // equilibrium flux: c (u + V) plus the anti-trapping current
            //   j_at = -a_t W (1 + (1-k) U) dphi/dt grad phi / |grad phi|
            R jx = C * (ux + vfac * gphi[0]), jy = C * (uy + vfac * gphi[1]),
              jz = C * (uz + vfac * gphi[2]);
            if constexpr (PQ > 0) {
                const R gg = gphi[0] * gphi[0] + gphi[1] * gphi[1] + gphi[2] * gphi[2];
                if (a.at_w != 0.0 && gg > (R)1e-20) {
                    const R jat = -(R)a.at_w * ((R)1 + ((R)1 - (R)a.k_part) * U) * dphi / sqrt(gg);
                    jx += jat * gphi[0];
                    jy += jat * gphi[1];
                    jz += jat * gphi[2];
                }
            }
            const R src = (R)a.latent_C * dphi;
#pragma unroll
            for (int i = 0; i < 7; i++) {
                const R w = (R)wgt<7>(i);
                const R cj = cx(i) * jx + cy(i) * jy + cz(i) * jz;
                // deviation form: c~eq = w (c - ref) + w c_i.j / cs^2
                const R ceq = w * (sc + cj / (R)0.25);
                cc[i] = cc[i] - om * (cc[i] - ceq) + w * src;
            }
            FaceState<R> fs;
            fs.dscale = A;  // Dirichlet faces fix U: c_w = A(phi) U_w
            push<7, ODD, Z, R, ZC>(cc, a.c, S, c, a, kKindScalar, zwc, zmc, tid, act, fs);
        }
    }

    // ---- phase field collide-stream (mkPhaseField::propagate +
    //      relaxation_anisotropy_thermal / _solute)
    if constexpr (PQ > 0) {
        if (a.g.adv) {
            if constexpr (!SO) load_phase_stencil<R>(a, c, P);
            R upf = 0;
            if (a.pf_coupling == 1) upf = T;
            if (a.pf_coupling == 2) upf = U;
            R Mrot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            bool rot = false;
            if (a.ngrains > 0) {
                const int gid = a.grain[c.id];
                if (gid > 0) {
#pragma unroll
                    for (int k = 0; k < 9; k++) Mrot[k] = (R)a.rot[(gid - 1) * 9 + k];
                    rot = true;
                }
            }
            R as;
            const Aniso<R> an{a.aniso_kind, (R)a.eps4, (R)a.aniso[0], (R)a.aniso[1]};
            const R Sg = phase_source<R>(P, upf, an, (R)a.lambda, (R)a.pf_dt,
                                         (R)a.pf_wdx2, Mrot, rot, as);
            // the populations are read after the source so that the stencil values are dead
            R g[PQ];
            pull<PQ, ODD, Z, R, ZN>(g, a.g, S, c, a, kKindScalar, zrf, tid);
            const R om = (R)a.omega_g;
            const R dp = phi - (R)a.g.ref;
            // advective flux (phi + 1) u: the liquid is invariant under any melt velocity
            const R pa = (phi + (R)1) / (R)cs2<PQ>();
#pragma unroll
            for (int i = 0; i < PQ; i++) {
                const R w = (R)wgt<PQ>(i);
                const R cu = cx(i) * ux + cy(i) * uy + cz(i) * uz;
                const R geq = w * (dp + pa * cu);
                g[i] = g[i] - om * (g[i] - geq) + w * Sg;
            }
            push<PQ, ODD, Z, R, ZN>(g, a.g, S, c, a, kKindScalar, zwf, zmf, tid, act,
                                    FaceState<R>());
        }
    }
}

// Host-side launch helpers for one storage type.
template <class R, int FQ, int PQ, bool TH, bool SO>
inline void run_step(int odd, const StepArgs& a, hipStream_t s) {
    dim3 grid(a.nblk), block(256);
    if (odd && a.zstage)
        k_step<R, FQ, PQ, TH, SO, true, true><<<grid, block, 0, s>>>(a);
    else if (odd)
        k_step<R, FQ, PQ, TH, SO, true><<<grid, block, 0, s>>>(a);
    else
        k_step<R, FQ, PQ, TH, SO, false><<<grid, block, 0, s>>>(a);
}

template <class R, int PQ>
inline void run_phase_moments(int odd, const StepArgs& a, void* phi_out, hipStream_t s) {
    dim3 grid(a.nblk), block(256);
    if (odd)
        k_phase_moments<R, PQ, true><<<grid, block, 0, s>>>(a, (R*)phi_out);
    else
        k_phase_moments<R, PQ, false><<<grid, block, 0, s>>>(a, (R*)phi_out);
}

// Lattice combinations with an instantiation: (flow Q, phase Q, thermal, solute).
#define MKG_COMBOS(X)                                                                       \
    X(19, 0, 0, 0) X(27, 0, 0, 0) X(19, 0, 1, 0) X(27, 0, 1, 0) X(0, 19, 1, 0) X(0, 7, 1, 0) \
    X(19, 19, 1, 0) X(19, 7, 1, 0) X(27, 19, 1, 0) X(19, 19, 0, 1) X(19, 7, 0, 1)           \
    X(0, 19, 0, 1) X(0, 0, 1, 0) X(0, 0, 0, 1) X(0, 19, 0, 0) X(0, 7, 0, 0)

}  // namespace mkg
