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
#include <type_traits>

#include "device.h"

namespace mkg {

template <class R>
__device__ inline R clamp01(R v) {
    return v < (R)0 ? (R)0 : (v > (R)1 ? (R)1 : v);
}

// 19-point phase stencil around the cell (periodic faces wrap, other faces replicate the
// edge value), index (dx+1)*9 + (dy+1)*3 + (dz+1).
template <class R, bool FACES = true, uint32_t FM = kAllFaces>
__device__ inline void load_phase_stencil(const StepArgs& a, const Cell& c, R (&P)[27]) {
    const R* __restrict__ ph = (const R*)a.phi_cur;
    const Nbh o = replicated_nbh(c, FACES ? (uint32_t)a.g.npmask & c.bits & FM : 0u);
#pragma unroll
    for (int i = 0; i < 19; i++)
        P[(cx(i) + 1) * 9 + (cy(i) + 1) * 3 + (cz(i) + 1)] = *slot_at(ph, o.at(cx(i), cy(i), cz(i)));
}

// Velocity faces (non-equilibrium extrapolation, mkFluidsUser::boundary_vel,
// MaterKit/src/mkFluidsUser.cpp:455-505): a face cell of a velocity face replaces its
// post-collision populations by the ones the pre-pass k_nee formed from its inward neighbour
// (unless that neighbour is a wall cell), before they stream.
template <int Q, class R>
__device__ inline void nee_replace(R* f, const StepArgs& a, const Cell& c) {
    const uint32_t hit = c.bits & (uint32_t)a.nee_mask;
    if (!hit) return;
    const int fc = __builtin_ctz(hit);
    // the inward neighbour, from constant offsets only: a runtime index into the cell's
    // offset arrays would move the whole Cell record to scratch memory
    const uint32_t n = fc == 0   ? nbr(c, 1, 0, 0)
                       : fc == 1 ? nbr(c, -1, 0, 0)
                       : fc == 2 ? nbr(c, 0, 1, 0)
                       : fc == 3 ? nbr(c, 0, -1, 0)
                       : fc == 4 ? nbr(c, 0, 0, 1)
                                 : nbr(c, 0, 0, -1);
    if (a.flag[n] == kIsWalls) return;
    const R* __restrict__ nb = (const R*)a.nee;
    const uint32_t k = face_slot(a, c, fc);
#pragma unroll
    for (int i = 0; i < Q; i++) f[i] = nb[(size_t)i * a.F + k];
}

// Pre-pass of a step with velocity faces, one thread per face-buffer index: for a cell b of
// a velocity face and its inward neighbour n, the collision of n exactly as k_step forms it
// (moments, body force with buoyancy, solid fraction, generalised moment relaxation), then
// f*_b = f*_n + f_eq(rho_n, u_b) - f_eq(rho_n, u_n) with u_n = sum c f*_n / rho_n (:485-500)
// and u_b the face velocity; the equilibrium is the case's (compressible or He-Luo).
template <class R, int FQ, bool TH, bool PH>
__global__ void __launch_bounds__(256) k_nee(const StepArgs a) {
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= a.F) return;
    const uint32_t nx = (uint32_t)a.nx, ny = (uint32_t)a.ny, nz = (uint32_t)a.nz;
    const uint32_t A = ny * nz, B = nx * nz, C = nx * ny;
    int fc;
    uint32_t x, y, z;
    if (k < 2u * A) {  // the inverse of face_slot
        fc = (int)(k / A);
        const uint32_t r = k % A;
        y = r / nz;
        z = r % nz;
        x = (fc & 1) ? nx - 1u : 0u;
    } else if (k < 2u * (A + B)) {
        const uint32_t q = k - 2u * A;
        fc = 2 + (int)(q / B);
        const uint32_t r = q % B;
        x = r / nz;
        z = r % nz;
        y = (fc & 1) ? ny - 1u : 0u;
    } else {
        const uint32_t q = k - 2u * (A + B);
        fc = 4 + (int)(q / C);
        const uint32_t r = q % C;
        x = r / ny;
        y = r % ny;
        z = (fc & 1) ? nz - 1u : 0u;
    }
    if (!((a.nee_mask >> fc) & 1)) return;
    Cell cb, cn;
    locate_at(a, (x * ny + y) * nz + z, cb);
    const int ax = fc >> 1, sg = (fc & 1) ? -1 : 1;
    locate_at(a, nbr(cb, ax == 0 ? sg : 0, ax == 1 ? sg : 0, ax == 2 ? sg : 0), cn);
    const int8_t fl = a.flag[cn.id];
    if (a.flag[cb.id] == kIsWalls || fl == kIsWalls) return;
    R ns = 0;
    if constexpr (PH) {
        const R phi = *slot_at((const R*)a.phi_cur, cn.id);
        if (fl == kIsSolid)
            ns = (R)1;
        else if (fl == kIsInter)
            ns = clamp01((R)0.5 * ((R)1 + phi));
    }
    R T = 0;
    if constexpr (TH) {
        R h[7];
        read_pops_rt<7, R>(h, a.h, cn, a, kKindScalar);
        R sh = 0;
#pragma unroll
        for (int i = 0; i < 7; i++) sh += h[i];
        T = (R)a.h.ref + sh;
    }
    R f[FQ];
    read_pops_rt<FQ, R>(f, a.f, cn, a, kKindFlow);
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
    const R den = a.incomp ? (R)a.rho0 : rho;
    const R ux = (jx + (R)0.5 * Fx) / den, uy = (jy + (R)0.5 * Fy) / den,
            uz = (jz + (R)0.5 * Fz) / den;
    flow_collide<FQ, R>(f, dr, den, ux, uy, uz, Fx, Fy, Fz, (R)a.s_nu, (R)a.s_bulk, (R)a.s_odd,
                        (R)a.s_even, ns);
    // the extrapolation: rho_n and u_n of the post-collision populations
    R pr = 0, px = 0, py = 0, pz = 0;
#pragma unroll
    for (int i = 0; i < FQ; i++) {
        pr += f[i];
        if (cx(i)) px += (R)cx(i) * f[i];
        if (cy(i)) py += (R)cy(i) * f[i];
        if (cz(i)) pz += (R)cz(i) * f[i];
    }
    const R rn = (R)a.f.ref + pr;
    const R vx = px / rn, vy = py / rn, vz = pz / rn;
    // face velocity (selected without a runtime index into the argument block)
    double ub[3] = {0, 0, 0};
#pragma unroll
    for (int f6 = 0; f6 < 6; f6++)
        if (fc == f6)
            for (int d = 0; d < 3; d++) ub[d] = a.u_face[f6][d];
    if (a.wall_rot) {
        double t[3];
        wall_rotation(a, fc, cb, t);
        for (int d = 0; d < 3; d++) ub[d] += t[d];
    }
    const R bx = (R)ub[0], by = (R)ub[1], bz = (R)ub[2];
    const R de = a.incomp ? (R)a.rho0 : rn;  // density of the velocity terms
    const R uu_b = bx * bx + by * by + bz * bz, uu_n = vx * vx + vy * vy + vz * vz;
    R* __restrict__ out = (R*)a.nee;
#pragma unroll
    for (int i = 0; i < FQ; i++) {
        const R cb_ = cx(i) * bx + cy(i) * by + cz(i) * bz;
        const R cn_ = cx(i) * vx + cy(i) * vy + cz(i) * vz;
        const R d = (R)wgt<FQ>(i) * de *
                    ((R)3 * (cb_ - cn_) + (R)4.5 * (cb_ * cb_ - cn_ * cn_) - (R)1.5 * (uu_b - uu_n));
        out[(size_t)i * a.F + k] = f[i] + d;
    }
}

// ZS: O step with the z stage (the block holds whole z rows): the c_z != 0 populations are
// loaded at the cell's own z and shifted through LDS, as in k_step.  PER: every face periodic
// and no wall cells (no face rules).  FM (with ZS): the faces the face rules check; ZLO | ZHI
// when the phase lattice's only non-periodic faces are z faces and there are no wall cells.
template <class R, int PQ, bool ODD, bool ZS = false, bool PER = false, uint32_t FM = kAllFaces>
__global__ void __launch_bounds__(256) k_phase_moments(const StepArgs a, R* __restrict__ phi_out) {
    constexpr bool Z = ODD && ZS;
    __shared__ R zv[Z ? nzdirs<PQ>() : 1][256];
    Cell c;
    bool act = true;
    if constexpr (Z) {
        act = locate_clamped(a, c);
    } else {
        if (!locate(a, c)) return;
    }
    const int8_t fl = a.flag[c.id];
    R g[PQ];
    if constexpr (Z) {
        const uint32_t own = zpull_issue<PQ, R, !PER, FM>(g, a.g, c, a, zv, threadIdx.x);
        __syncthreads();
        zpull_finish<PQ, R>(g, c, a, zv, threadIdx.x, own);
        if (!act) return;
    }
    if (fl == kIsWalls) {
        // wall cells keep their phase value: carry it into the buffer being written
        phi_out[c.id] = ((const R*)a.phi_prev)[c.id];
        return;
    }
    if constexpr (!Z) read_pops<PQ, ODD, R, !PER>(g, a.g, c, a, kKindScalar);
    R s = 0;
#pragma unroll
    for (int i = 0; i < PQ; i++) s += g[i];
    const R phi = (R)a.g.ref + s;
    phi_out[c.id] = phi;
    const int8_t nf = classify(fl, (double)phi, a.phi_s, a.pf_model);
    if (nf != fl) a.flag[c.id] = nf;
}

// Minimum waves per SIMD requested from the register allocator for the fused kernel.  FP32:
// four (at most 128 VGPRs).  Left alone, the allocator gives the z-staged O step holding both
// bodies 136 VGPRs and three waves; with four it spills (72 bytes of scratch per lane) and the
// step runs 7 % faster.  The all-periodic variant (PER) needs 115 VGPRs and does not spill.
// FP64 kernels need about 230 VGPRs (two waves); a bound there only adds spills.
#ifndef MKG_STEP_WAVES_F32
#define MKG_STEP_WAVES_F32 4
#endif
#ifndef MKG_STEP_WAVES_F32_ZFACES
#define MKG_STEP_WAVES_F32_ZFACES MKG_STEP_WAVES_F32
#endif

// LDS slots of the z stage: flow pull region (reused by the phase push), flow push region,
// phase pull region.
template <int FQ, int PQ>
constexpr int zstage_slots() {
    int n = 1;
    if (FQ == 19 && nzdirs<19>() > n) n = nzdirs<19>();
    if (FQ == 27 && nzdirs<27>() > n) n = nzdirs<27>();
    if (PQ == 19 && nzdirs<19>() > n) n = nzdirs<19>();
    if (PQ == 7 && nzdirs<7>() > n) n = nzdirs<7>();
    return n;
}
template <int Q>
constexpr int zslots_of() {
    return Q == 19 ? nzdirs<19>() : (Q == 27 ? nzdirs<27>() : (Q == 7 ? nzdirs<7>() : 1));
}

// PER: every face of every lattice periodic, no wall cells, and (with ZS) no thread past the
// last cell: the kernel holds only the body without face rules, which the register allocator
// then fits in 128 VGPRs with fewer spills than the kernel holding both bodies.
// ZS: O step with the z stage (only instantiated with ODD; the block must hold whole z rows).
// With it every thread stays to the end (barriers), and act marks the threads that update.
// The z stage runs in three block barriers:
//   B1  after the pull loads of the thermal, solute and flow lattices
//   B2  after the flow / thermal / solute push stages and the phase pull loads (which are
//       issued after the phase source, so that the stencil values are dead by then)
//   B3  after the phase push stage, which reuses the flow pull region (last read before B2)
// Every other stage region is written once and read once, on either side of one barrier.
// TAPE: the step also writes the adjoint tape: phi^n and the pre-collision thermal
// populations of every cell (the phase-field + thermal model, with or without the D3Q19
// flow), or the pre-collision flow populations (the flow alone), slot i at tape + i*S.
template <class R, int FQ, int PQ, bool TH, bool SO, bool ODD, bool ZS = false, bool PER = false,
          bool TAPE = false, uint32_t SFM = kAllFaces>
__global__ void __launch_bounds__(256, sizeof(R) == 4 ? ((ODD && ZS && !PER)
                                                             ? MKG_STEP_WAVES_F32_ZFACES
                                                             : MKG_STEP_WAVES_F32)
                                                      : 1)
    k_step(const StepArgs a) {
    constexpr bool Z = ODD && ZS;
    constexpr int ZF = Z ? zstage_slots<FQ, PQ>() : 1;
    constexpr int ZW = (Z && FQ > 0) ? zslots_of<FQ>() : 1;
    constexpr int ZG = (Z && PQ > 0) ? zslots_of<PQ>() : 1;
    constexpr int ZH = (Z && TH) ? nzdirs<7>() : 1;
    constexpr int ZC = (Z && SO) ? nzdirs<7>() : 1;
    __shared__ R zrf[ZF][256], zwf[ZW][256], zrg[ZG][256];  // flow pull (phase push), flow push, phase pull
    __shared__ R zrh[ZH][256], zwh[ZH][256];  // thermal
    __shared__ R zrc[ZC][256], zwc[ZC][256];  // solute
    __shared__ uint32_t zmf[(Z && FQ > 0) ? 256 : 1], zmg[(Z && PQ > 0) ? 256 : 1],
        zmh[(Z && TH) ? 256 : 1], zmc[(Z && SO) ? 256 : 1];
    const uint32_t tid = threadIdx.x;
    // Blocks (z stage) or waves (otherwise) none of whose cells lies on a non-periodic face,
    // in a domain without wall cells, take the body without face rules (FM = 0); a z stage
    // block must also lie wholly before the last cell, so that every thread updates.  The
    // others take the body with the face rules of SFM (every face, or only the z faces in a
    // domain whose only non-periodic faces are z faces, with no wall cells or velocity faces).
    constexpr bool ZROWS = Z && SFM == (ZLO | ZHI);
    bool slow = !PER && a.has_walls != 0;
    if (!PER && !ZROWS && !slow) {
        const uint32_t t = thread_cell(a);
        Cell e;
        locate_at(a, t < a.N ? t : a.N - 1u, e);
        const uint32_t np = (uint32_t)(a.f.npmask | a.g.npmask | a.h.npmask | a.c.npmask);
        const int face = (e.bits & np) != 0u;
        if constexpr (Z)
            slow = __syncthreads_or(face || t >= a.N) != 0;
        else
            slow = __ballot(face) != 0ull;
    }
    auto body = [&](auto faces_tag) {
    constexpr uint32_t FM = decltype(faces_tag)::value;
    constexpr bool FACES = FM != 0u;
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
        if constexpr (!FACES) act = true;  // no walls, no thread past the last cell
    } else {
        if (fl == kIsWalls) return;
    }
    const uint32_t S = a.S;

    R phi = (R)-1, dphi = (R)0, ns = (R)0;
    if constexpr (PQ > 0) {
        phi = *slot_at((const R*)a.phi_cur, c.id);
        dphi = phi - *slot_at((const R*)a.phi_prev, c.id);
        if (fl == kIsSolid)
            ns = (R)1;
        else if (fl == kIsInter)
            ns = clamp01((R)0.5 * ((R)1 + phi));
    }

    // ---- population loads at time level n: thermal, solute, flow
    R h[TH ? 7 : 1];
    R cc[SO ? 7 : 1];
    R f[FQ > 0 ? FQ : 1];
    uint32_t own_h = 0, own_c = 0, own_f = 0;
    if constexpr (TH) {
        if (!a.h.adv)
            read_pops_rt<7, R>(h, a.h, c, a, kKindScalar);
        else if constexpr (Z)
            own_h = zpull_issue<7, R, FACES, FM>(h, a.h, c, a, zrh, tid);
        else
            read_pops<7, ODD, R, FACES, FM>(h, a.h, c, a, kKindScalar);
    }
    if constexpr (SO) {
        if (!a.c.adv)
            read_pops_rt<7, R>(cc, a.c, c, a, kKindScalar);
        else if constexpr (Z)
            own_c = zpull_issue<7, R, FACES, FM>(cc, a.c, c, a, zrc, tid);
        else
            read_pops<7, ODD, R, FACES, FM>(cc, a.c, c, a, kKindScalar);
    }
    if constexpr (FQ > 0) {
        if (!a.f.adv)
            read_pops_rt<FQ, R>(f, a.f, c, a, kKindFlow);
        else if constexpr (Z)
            own_f = zpull_issue<FQ, R, FACES, FM>(f, a.f, c, a, zrf, tid);
        else
            read_pops<FQ, ODD, R, FACES, FM>(f, a.f, c, a, kKindFlow);
    }
    if constexpr (Z) {
        __syncthreads();  // B1
        if constexpr (TH)
            if (a.h.adv) zpull_finish<7, R>(h, c, a, zrh, tid, own_h);
        if constexpr (SO)
            if (a.c.adv) zpull_finish<7, R>(cc, c, a, zrc, tid, own_c);
        if constexpr (FQ > 0)
            if (a.f.adv) zpull_finish<FQ, R>(f, c, a, zrf, tid, own_f);
    }

    // ---- scalar moments at time level n
    R sh = 0, T = 0;
    if constexpr (TH) {
#pragma unroll
        for (int i = 0; i < 7; i++) sh += h[i];
        T = (R)a.h.ref + sh;
    }
    R sc = 0, C = 0, U = 0;
    if constexpr (SO) {
#pragma unroll
        for (int i = 0; i < 7; i++) sc += cc[i];
        C = (R)a.c.ref + sc;
        const R A = (R)0.5 * (((R)1 + (R)a.k_part) - ((R)1 - (R)a.k_part) * phi);
        U = C / A;
    }

    // ---- flow
    R ux = 0, uy = 0, uz = 0;
    if constexpr (FQ > 0) {
        if constexpr (TAPE && PQ == 0) {
            // the flow adjoint's tape: the pre-collision populations (deviations) of level n
            if (act) {
                R* __restrict__ tp = (R*)a.tape;
#pragma unroll
                for (int i = 0; i < FQ; i++) stp(slot_at(tp + (size_t)i * S, c.id), f[i]);
            }
        }
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
            if constexpr (FM == kAllFaces) nee_replace<FQ, R>(f, a, c);
            FaceState<R> fs;
            fs.wscale = (R)1 - ns;
            fs.ux = ux;
            fs.uy = uy;
            fs.uz = uz;
            if constexpr (Z)
                zpush_stage<FQ, R, FACES, FM>(f, a.f, c, a, kKindFlow, zwf, zmf, tid, act, fs);
            else
                write_pops<FQ, ODD, R, FACES, FM>(f, a.f, c, a, kKindFlow, fs);
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
            if constexpr (TAPE) {
                if (act) {
                    R* __restrict__ tp = (R*)a.tape;
                    stp(slot_at(tp, c.id), phi);
#pragma unroll
                    for (int i = 0; i < 7; i++) stp(slot_at(tp + (size_t)(1 + i) * S, c.id), h[i]);
                }
            }
#pragma unroll
            for (int i = 0; i < 7; i++) {
                const R w = (R)wgt<7>(i);
                const R cu = cx(i) * ux + cy(i) * uy + cz(i) * uz;
                const R heq = w * (sh + T * cu / (R)0.25);
                h[i] = h[i] - om * (h[i] - heq) + w * src;
            }
            if constexpr (Z)
                zpush_stage<7, R, FACES, FM>(h, a.h, c, a, kKindScalar, zwh, zmh, tid, act,
                                  FaceState<R>());
            else
                write_pops<7, ODD, R, FACES, FM>(h, a.h, c, a, kKindScalar);
        }
    }

    // ---- solute collide-stream (mkSolutalField::propagate + relaxation): the conserved
    //      c = A(phi) U obeys dc/dt + div(c u) = div(D q grad U), written as an
    //      advection-diffusion of c with D_c = D q / A and the drift V = D q grad A / A^2
    R P[27];
    if constexpr (SO && PQ > 0) load_phase_stencil<R, FACES, FM>(a, c, P);
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
            if constexpr (Z)
                zpush_stage<7, R, FACES, FM>(cc, a.c, c, a, kKindScalar, zwc, zmc, tid, act, fs);
            else
                write_pops<7, ODD, R, FACES, FM>(cc, a.c, c, a, kKindScalar, fs);
        }
    }

    // ---- phase field collide-stream (mkPhaseField::propagate +
    //      relaxation_anisotropy_thermal / _solute)
    R g[PQ > 0 ? PQ : 1];
    R Sg = 0;
    R acn0 = 0, acn1 = 0, acn2 = 0, acz = 0;  // Allen-Cahn normal and counter-term scale
    uint32_t own_g = 0;
    if constexpr (PQ > 0) {
        if (a.g.adv) {
            if constexpr (!SO) load_phase_stencil<R, FACES, FM>(a, c, P);
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
            if (a.pf_model == 1) {
                // conservative Allen-Cahn (relaxation_AllenCahn): normal of the LB-weighted
                // gradient s = sum_i c_i w_i phi(x + c_i), counter-term
                // (1 - omega/2) w_i (c_i . n)(1 - phi^2) dx / (sqrt 2 W0) of Chai et al.
                // (2018), growth source -lambda u (1 - phi^2)^2 dt
                R gr[3];
                stencil_gradient<R>(P, gr);
                const R sn = sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]) +
                             (R)(3.0 * 1e-12);
                const R d = (R)1 - phi * phi;
                acn0 = gr[0] / sn;
                acn1 = gr[1] / sn;
                acn2 = gr[2] / sn;
                acz = (R)a.ac_zc * d;
                Sg = -(R)a.lambda * upf * d * d * (R)a.pf_dt;
            } else {
                R as;
                const Aniso<R> an{a.aniso_kind, (R)a.eps4, (R)a.aniso[0], (R)a.aniso[1]};
                Sg = phase_source<R>(P, upf, an, (R)a.lambda, (R)a.pf_dt, (R)a.pf_wdx2, Mrot,
                                     rot, as);
            }
            // the populations are read after the source so that the stencil values are dead
            // (loaded with the other lattices, they measured 13 % slower at 256^3)
            if constexpr (Z)
                own_g = zpull_issue<PQ, R, FACES, FM>(g, a.g, c, a, zrg, tid);
            else
                read_pops<PQ, ODD, R, FACES, FM>(g, a.g, c, a, kKindScalar);
        }
    }
    if constexpr (Z) {
        __syncthreads();  // B2
        if constexpr (FQ > 0)
            if (a.f.adv) zpush_finish<FQ, R, FACES, FM>(a.f, c, a, zwf, zmf, tid, act);
        if constexpr (TH)
            if (a.h.adv) zpush_finish<7, R, FACES, FM>(a.h, c, a, zwh, zmh, tid, act);
        if constexpr (SO)
            if (a.c.adv) zpush_finish<7, R, FACES, FM>(a.c, c, a, zwc, zmc, tid, act);
        if constexpr (PQ > 0)
            if (a.g.adv) zpull_finish<PQ, R>(g, c, a, zrg, tid, own_g);
    }
    if constexpr (PQ > 0) {
        if (a.g.adv) {
            const R om = (R)a.omega_g;
            const R dp = phi - (R)a.g.ref;
            // advective flux (phi + 1) u: the liquid is invariant under any melt velocity
            const R pa = (phi + (R)1) / (R)cs2<PQ>();
#pragma unroll
            for (int i = 0; i < PQ; i++) {
                const R w = (R)wgt<PQ>(i);
                const R cu = cx(i) * ux + cy(i) * uy + cz(i) * uz;
                const R cn = cx(i) * acn0 + cy(i) * acn1 + cz(i) * acn2;
                const R geq = w * (dp + pa * cu);
                g[i] = g[i] - om * (g[i] - geq) + w * (Sg + acz * cn);
            }
            if constexpr (Z)
                zpush_stage<PQ, R, FACES, FM>(g, a.g, c, a, kKindScalar, zrf, zmg, tid, act,
                                   FaceState<R>());
            else
                write_pops<PQ, ODD, R, FACES, FM>(g, a.g, c, a, kKindScalar);
        }
        if constexpr (Z) {
            __syncthreads();  // B3
            if (a.g.adv) zpush_finish<PQ, R, FACES, FM>(a.g, c, a, zrf, zmg, tid, act);
        }
    }
    };
    // With the z stage and only z faces, every z row holds a face cell, so every block takes
    // the face body: the kernel holds it alone.
    if constexpr (PER)
        body(std::integral_constant<uint32_t, 0u>{});
    else if (ZROWS)
        body(std::integral_constant<uint32_t, SFM>{});
    else if (slow)
        body(std::integral_constant<uint32_t, SFM>{});
    else
        body(std::integral_constant<uint32_t, 0u>{});
}

// Host-side launch helpers for one storage type.
template <class R, int FQ, int PQ, bool TH, bool SO>
inline void run_step(int odd, const StepArgs& a, hipStream_t s) {
    dim3 grid(a.nblk), block(256);
    const bool per = !a.has_walls && (a.f.npmask | a.g.npmask | a.h.npmask | a.c.npmask) == 0;
    // the adjoint's forward steps (the phase-field + thermal model with or without the flow,
    // or the flow alone), with the kernel variant a plain step would take, so that a recompute
    // repeats the forward exactly
    if constexpr (((FQ == 0 || FQ == 19) && PQ == 19 && TH && !SO) ||
                  (FQ > 0 && PQ == 0 && !TH && !SO)) {
        if (a.tape) {
            if (odd && a.zstage) {
                if (per && a.N == a.nblk * 256u)
                    k_step<R, FQ, PQ, TH, SO, true, true, true, true><<<grid, block, 0, s>>>(a);
                else
                    k_step<R, FQ, PQ, TH, SO, true, true, false, true><<<grid, block, 0, s>>>(a);
            } else if (odd) {
                if (per)
                    k_step<R, FQ, PQ, TH, SO, true, false, true, true><<<grid, block, 0, s>>>(a);
                else
                    k_step<R, FQ, PQ, TH, SO, true, false, false, true><<<grid, block, 0, s>>>(a);
            } else {
                if (per)
                    k_step<R, FQ, PQ, TH, SO, false, false, true, true><<<grid, block, 0, s>>>(a);
                else
                    k_step<R, FQ, PQ, TH, SO, false, false, false, true><<<grid, block, 0, s>>>(a);
            }
            return;
        }
    }
    // only z faces non-periodic, no wall cells, no velocity faces: the face body checks the
    // z faces alone
    const uint32_t np = (uint32_t)(a.f.npmask | a.g.npmask | a.h.npmask | a.c.npmask);
    const bool zonly = !per && !a.has_walls && !a.nee_mask && (np & ~(uint32_t)(ZLO | ZHI)) == 0u;
    constexpr uint32_t ZF = ZLO | ZHI;
    if (odd && a.zstage) {
        if (per && a.N == a.nblk * 256u)
            k_step<R, FQ, PQ, TH, SO, true, true, true><<<grid, block, 0, s>>>(a);
        else if (zonly)
            k_step<R, FQ, PQ, TH, SO, true, true, false, false, ZF><<<grid, block, 0, s>>>(a);
        else
            k_step<R, FQ, PQ, TH, SO, true, true><<<grid, block, 0, s>>>(a);
    } else if (odd) {
        if (per)
            k_step<R, FQ, PQ, TH, SO, true, false, true><<<grid, block, 0, s>>>(a);
        else if (zonly)
            k_step<R, FQ, PQ, TH, SO, true, false, false, false, ZF><<<grid, block, 0, s>>>(a);
        else
            k_step<R, FQ, PQ, TH, SO, true><<<grid, block, 0, s>>>(a);
    } else {
        if (per)
            k_step<R, FQ, PQ, TH, SO, false, false, true><<<grid, block, 0, s>>>(a);
        else if (zonly)
            k_step<R, FQ, PQ, TH, SO, false, false, false, false, ZF><<<grid, block, 0, s>>>(a);
        else
            k_step<R, FQ, PQ, TH, SO, false><<<grid, block, 0, s>>>(a);
    }
}

template <class R, int FQ, int PQ, bool TH>
inline void run_nee(const StepArgs& a, hipStream_t s) {
    if constexpr (FQ > 0) {
        dim3 grid((a.F + 255u) / 256u), block(256);
        k_nee<R, FQ, TH, (PQ > 0)><<<grid, block, 0, s>>>(a);
    }
}

template <class R, int PQ>
inline void run_phase_moments(int odd, const StepArgs& a, void* phi_out, hipStream_t s) {
    dim3 grid(a.nblk), block(256);
    const bool per = !a.has_walls && a.g.npmask == 0;
    const bool zonly = !a.has_walls && ((uint32_t)a.g.npmask & ~(uint32_t)(ZLO | ZHI)) == 0u;
    if (odd && a.zstage) {
        if (per)
            k_phase_moments<R, PQ, true, true, true><<<grid, block, 0, s>>>(a, (R*)phi_out);
        else if (zonly)
            k_phase_moments<R, PQ, true, true, false, ZLO | ZHI><<<grid, block, 0, s>>>(a, (R*)phi_out);
        else
            k_phase_moments<R, PQ, true, true><<<grid, block, 0, s>>>(a, (R*)phi_out);
    } else if (odd) {
        if (per)
            k_phase_moments<R, PQ, true, false, true><<<grid, block, 0, s>>>(a, (R*)phi_out);
        else
            k_phase_moments<R, PQ, true><<<grid, block, 0, s>>>(a, (R*)phi_out);
    } else {
        k_phase_moments<R, PQ, false><<<grid, block, 0, s>>>(a, (R*)phi_out);
    }
}

// Lattice combinations with an instantiation: (flow Q, phase Q, thermal, solute).
#define MKG_COMBOS(X)                                                                       \
    X(19, 0, 0, 0) X(27, 0, 0, 0) X(19, 0, 1, 0) X(27, 0, 1, 0) X(0, 19, 1, 0) X(0, 7, 1, 0) \
    X(19, 19, 1, 0) X(19, 7, 1, 0) X(27, 19, 1, 0) X(19, 19, 0, 1) X(19, 7, 0, 1)           \
    X(0, 19, 0, 1) X(0, 0, 1, 0) X(0, 0, 0, 1) X(0, 19, 0, 0) X(0, 7, 0, 0) X(19, 19, 0, 0)

}  // namespace mkg
