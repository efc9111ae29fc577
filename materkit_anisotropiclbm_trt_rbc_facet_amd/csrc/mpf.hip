// Multi-phase field: n order parameters phi_j in [-1, 1], each with its own D3Q19 population
// lattice, plus an optional D3Q7 thermal lattice -- the per-grain lattices gVec / phiVec of
// mkFluidsUser::propagate / relaxation (MaterKit/src/mkFluidsUser.cpp:68-134, 136-299) as run
// by run_multiPhaseField (MaterKit/MaterKit.cpp:529-576).  Same AA-pattern storage, face rules
// and wall handling as the fused solver step (device.h).
//
// One step is n + 1 launches:
//   k_mpf_moments  phi_j = sum_i g_j,i for every cell and the combined phi = (n-1) + sum phi_j
//                  (the first sweep of mkFluidsUser::relaxation, :147-172); the gradient
//                  stencils of the collision read the neighbours' phi_j
//   k_mpf_one      (one launch per order parameter j) per cell: LB-weighted gradient
//                  s_j = sum_i c_i w_i phi_j(x+c_i)
//                  and normal n_j = s_j / (|s_j| + eps) (:208-217), equilibrium with phi_j as the
//                  density and the order parameter's own velocity (getEquPDFs, :219-230), the
//                  interaction P_j = -2 kappa_j (phi_j/2 + 1/2) sum_{k != j} (phi_k/2 + 1/2)^2
//                  (:236-256), the growth term Q_j (:266-271), the conservative Allen-Cahn
//                  counter-term (1 - omega/2) w_i (c_i . n_j) (1 - phi_j^2) dx / (sqrt 2 W0)
//                  (getAddedPhasePDFs, Chai et al. 2018, :273-279), BGK with
//                  omega_g = 1/(dt/dx^2/cs^2 + 1/2) (:282-293), then streaming; the launch of
//                  the last order parameter also collides and streams the thermal lattice
//                  (latent source from the combined dphi).
// Cells are not re-classified: a cell other than fluid does not collide in the template
// (:185-188), so classifying interface cells would freeze every grain.  Flag 3 (walls) is the
// only non-fluid code used here.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/mkg.h"
#include "device.h"

namespace mkg {

void set_error(const std::string& msg);

struct MpfArgs {
    StepArgs a;    // geometry, flags and walls; a.g = lattice of order parameter 0, a.h thermal
    LatArgs gl[MKG_MPF_MAX];  // lattice of each order parameter (a.g with its own base)
    int32_t n, thermal;
    double u[MKG_MPF_MAX][3];
    double kappa[MKG_MPF_MAX];
    double growth, lambda, eps;
    double omega_g;  // 1 / (dt / wdx^-2 / cs^2 + 1/2)
    double zc;       // (1 - omega_g / 2) / (sqrt 2 W0/dx): scale of the Allen-Cahn counter-term
    double dt, omega_h, latent_T, T_ref;
    // Karma-Rappel grains and the solute lattice
    int32_t coupling;  // 0 none, 1 T, 2 U
    int32_t need_T, need_U;  // launches other than the last read T / U from the populations
    int32_t akind[MKG_MPF_MAX];
    int32_t rotated[MKG_MPF_MAX];  // 0: identity rotation
    double e4[MKG_MPF_MAX], ap[MKG_MPF_MAX][2], rot[MKG_MPF_MAX][9];
    double wdx2;
    double D_l, D_s, k_part, q_min, latent_C, at_w;
};

// ZS (with ODD): the block holds whole z rows, and the z-moving populations of every order
// parameter are loaded at the cell's own z and shifted through LDS as in k_phase_moments (two
// LDS regions in turn, one barrier per order parameter; every thread takes part).  PER: no
// face rules (every face periodic, no wall cells).
template <class R, bool ODD, bool ZS = false, bool PER = false>
__global__ void __launch_bounds__(256) k_mpf_moments(const MpfArgs m, R* __restrict__ out) {
    constexpr bool Z = ODD && ZS;
    __shared__ R zv[Z ? 2 : 1][Z ? nzdirs<19>() : 1][256];
    Cell c;
    bool act = true;
    if constexpr (Z) {
        act = locate_clamped(m.a, c);
    } else {
        if (!locate(m.a, c)) return;
    }
    const uint32_t S = m.a.S;
    const R* __restrict__ prev = (const R*)m.a.phi_prev;
    const bool wall = m.a.flag[c.id] == kIsWalls;
    if (!Z && wall) {
        for (int j = 0; j <= m.n; j++)
            *slot_at(out + (size_t)j * S, c.id) = *slot_at(prev + (size_t)j * S, c.id);
        return;
    }
    const bool put = act && !wall;
    R comb = (R)(m.n - 1);
#pragma unroll
    for (int j = 0; j < MKG_MPF_MAX; j++) {
        if (j < m.n) {
            const LatArgs& L = m.gl[j];  // j is unrolled: a fixed offset, no copy
            R g[19];
            if constexpr (Z) {
                const uint32_t own = zpull_issue<19, R, !PER>(g, L, c, m.a, zv[j & 1], threadIdx.x);
                __syncthreads();
                zpull_finish<19, R>(g, c, m.a, zv[j & 1], threadIdx.x, own);
            } else {
                read_pops<19, ODD, R, !PER>(g, L, c, m.a, kKindScalar);
            }
            R s = 0;
#pragma unroll
            for (int i = 0; i < 19; i++) s += g[i];
            const R ph = (R)L.ref + s;
            if (put) *slot_at(out + (size_t)j * S, c.id) = ph;
            comb += ph;
        }
    }
    if (!act) return;
    if (wall) {
        // wall cells keep their order parameters: carry them into the buffer being written
        for (int j = 0; j <= m.n; j++)
            *slot_at(out + (size_t)j * S, c.id) = *slot_at(prev + (size_t)j * S, c.id);
        return;
    }
    *slot_at(out + (size_t)m.n * S, c.id) = comb;
}

// LB-weighted gradient of order parameter plane ph around the cell (no 1/cs^2, as in the
// template): periodic faces wrap, other faces replicate the edge value.
template <class R>
__device__ inline void mpf_gradient(const StepArgs& a, const Cell& c, const R* ph, R (&s)[3]) {
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
    s[0] = s[1] = s[2] = 0;
#pragma unroll
    for (int i = 1; i < 19; i++) {
        const R v = (R)wgt<19>(i) * *slot_at(ph, px[cx(i) + 1] + py[cy(i) + 1] + pz[cz(i) + 1]);
        if (cx(i)) s[0] += (R)cx(i) * v;
        if (cy(i)) s[1] += (R)cy(i) * v;
        if (cz(i)) s[2] += (R)cz(i) * v;
    }
}

template <class R>
__device__ inline R clamp_pm1(R v) {
    return v < (R)-1 ? (R)-1 : (v > (R)1 ? (R)1 : v);
}

// The 3 x 3 x 3 neighbourhood of order parameter plane ph around the cell (the 19 points of
// D3Q19; corners unused): periodic faces wrap, other faces replicate the edge value.
template <class R>
__device__ inline void mpf_stencil(const StepArgs& a, const Cell& c, const R* ph, R (&P)[27]) {
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
    for (int k = 0; k < 27; k++) P[k] = 0;
#pragma unroll
    for (int i = 0; i < 19; i++)
        P[(cx(i) + 1) * 9 + (cy(i) + 1) * 3 + (cz(i) + 1)] =
            *slot_at(ph, px[cx(i) + 1] + py[cy(i) + 1] + pz[cz(i) + 1]);
}

#ifndef MKG_MPF_KR_WAVES_F32
#define MKG_MPF_KR_WAVES_F32 4
#endif

// Collide-stream of one order parameter j (one launch per order parameter; the thermal and
// solute lattices ride on the last one).  The populations of order parameter j are updated in
// place, which no other launch of the step reads; T and U for the growth terms come from the
// thermal and solute populations, which only the last launch updates.  O steps take the z
// shift through LDS when a block holds whole z rows (ZS): on the last launch one barrier after
// the scalar pull loads, then one after the phase pull loads and one after the phase push
// stage.  The last launch collides the scalar lattices first, so that their populations are
// dead before the phase stencil is loaded.
// KR: Karma-Rappel grain (the single-grain phase source with the grain's anisotropy and
// rotation, interaction on the time scale tau0 a_s^2), else the template's conservative
// Allen-Cahn order parameter.
// Launched per order parameter, the kernel holds one lattice in registers (four waves per
// SIMD); the earlier single launch over all order parameters took 168-232 VGPRs.
template <class R, bool TH, bool SO, bool KR, bool ODD, bool ZS = false>
__global__ void __launch_bounds__(256, sizeof(R) == 4 ? (KR ? MKG_MPF_KR_WAVES_F32 : 4) : 1)
    k_mpf_one(const MpfArgs m, const LatArgs L, const int j) {
    constexpr bool Z = ODD && ZS;
    constexpr int ZQ = Z ? nzdirs<19>() : 1;
    constexpr int ZT = (Z && TH) ? nzdirs<7>() : 1;
    constexpr int ZC = (Z && SO) ? nzdirs<7>() : 1;
    __shared__ R zr[ZQ][256], zw[ZQ][256], zrh[ZT][256], zwh[ZT][256], zrc[ZC][256],
        zwc[ZC][256];
    __shared__ uint32_t zm[Z ? 256 : 1], zmh[(Z && TH) ? 256 : 1], zmc[(Z && SO) ? 256 : 1];
    const uint32_t tid = threadIdx.x;
    // Blocks (z stage) or waves none of whose cells lies on a non-periodic face, in a domain
    // without wall cells (and, with the z stage, wholly before the last cell), take the body
    // without face rules (FACES false), as the fused step does.
    bool slow = m.a.has_walls != 0;
    if (!slow) {
        const uint32_t t = thread_cell(m.a);
        Cell e;
        locate_at(m.a, t < m.a.N ? t : m.a.N - 1u, e);
        const uint32_t np = (uint32_t)(L.npmask | (TH ? m.a.h.npmask : 0) |
                                       (SO ? m.a.c.npmask : 0));
        const int face = (e.bits & np) != 0u;
        if constexpr (Z)
            slow = __syncthreads_or(face || t >= m.a.N) != 0;
        else
            slow = __ballot(face) != 0ull;
    }
    auto body = [&](auto faces_tag) {
    constexpr bool FACES = decltype(faces_tag)::value;
    Cell c;
    bool act = true;
    if constexpr (Z) {
        act = locate_clamped(m.a, c);
    } else {
        if (!locate(m.a, c)) return;
    }
    const bool wall = m.a.flag[c.id] == kIsWalls;
    if constexpr (Z) {
        act = act && !wall;
        if constexpr (!FACES) act = true;  // no walls, no thread past the last cell
    } else {
        if (wall) return;
    }
    const StepArgs& a = m.a;
    const uint32_t S = a.S;
    const R* __restrict__ cur = (const R*)a.phi_cur;
    const R* __restrict__ prev = (const R*)a.phi_prev;
    const R pj = *slot_at(cur + (size_t)j * S, c.id);
    // combined phi (moment pass of this step) and its change, for the scalar lattices
    const R pc = *slot_at(cur + (size_t)m.n * S, c.id);
    const R dpc = pc - *slot_at(prev + (size_t)m.n * S, c.id);

    // ---- scalar lattices (the last launch): pull loads, moments, collide, push stage
    R h[TH ? 7 : 1];
    R cc[SO ? 7 : 1];
    uint32_t own_h = 0, own_c = 0;
    if constexpr (TH) {
        if constexpr (Z)
            own_h = zpull_issue<7, R, FACES>(h, a.h, c, a, zrh, tid);
        else
            read_pops<7, ODD, R, FACES>(h, a.h, c, a, kKindScalar);
    }
    if constexpr (SO) {
        if constexpr (Z)
            own_c = zpull_issue<7, R, FACES>(cc, a.c, c, a, zrc, tid);
        else
            read_pops<7, ODD, R, FACES>(cc, a.c, c, a, kKindScalar);
    }
    if constexpr (Z && (TH || SO)) {
        __syncthreads();
        if constexpr (TH) zpull_finish<7, R>(h, c, a, zrh, tid, own_h);
        if constexpr (SO) zpull_finish<7, R>(cc, c, a, zrc, tid, own_c);
    }
    R T = (R)m.T_ref, U = 0;
    if constexpr (TH) {
        // thermal collide-stream (mkThermalField::propagate / relaxation of run_multiPhaseField)
        R sh = 0;
#pragma unroll
        for (int i = 0; i < 7; i++) sh += h[i];
        T = (R)a.h.ref + sh;
        const R omh = (R)m.omega_h, srh = (R)m.latent_T * dpc;
#pragma unroll
        for (int i = 0; i < 7; i++) {
            const R w = (R)wgt<7>(i);
            h[i] = h[i] - omh * (h[i] - w * sh) + w * srh;
        }
        if constexpr (Z)
            zpush_stage<7, R, FACES>(h, a.h, c, a, kKindScalar, zwh, zmh, tid, act,
                                     FaceState<R>());
        else
            write_pops<7, ODD, R, FACES>(h, a.h, c, a, kKindScalar);
    } else {
        if (m.need_T) {
            R hh[7];
            read_pops_rt<7, R>(hh, a.h, c, a, kKindScalar);
            R sh = 0;
#pragma unroll
            for (int i = 0; i < 7; i++) sh += hh[i];
            T = (R)a.h.ref + sh;
        }
    }
    // the solute sees the combined phi clamped to [-1, 1] (overlapping grains exceed 1, where
    // A(phi) would vanish)
    const R ps = clamp_pm1(pc);
    const R A = (R)0.5 * (((R)1 + (R)m.k_part) - ((R)1 - (R)m.k_part) * ps);
    if constexpr (SO) {
        // solute collide-stream: the single-grain solute equation with the clamped phi
        R sc = 0;
#pragma unroll
        for (int i = 0; i < 7; i++) sc += cc[i];
        const R C = (R)a.c.ref + sc;
        U = C / A;
        const R dps = ps - clamp_pm1(*slot_at(prev + (size_t)m.n * S, c.id));
        R Pc[27];
        mpf_stencil<R>(a, c, cur + (size_t)m.n * S, Pc);
#pragma unroll
        for (int k = 0; k < 27; k++) Pc[k] = clamp_pm1(Pc[k]);
        R gphi[3];
        stencil_gradient<R>(Pc, gphi);
        R q = (R)0.5 * ((R)1 - ps);
        if (m.D_l > 0) q += (R)(m.D_s / m.D_l) * (R)0.5 * ((R)1 + ps);
        if (q < (R)m.q_min) q = (R)m.q_min;
        const R Dc = (R)m.D_l * q / A;
        const R om = (R)1 / ((R)0.5 + Dc / (R)0.25);
        const R vfac = (R)(-0.5 * (1.0 - m.k_part)) * (R)m.D_l * q / (A * A);
        R jx = C * vfac * gphi[0], jy = C * vfac * gphi[1], jz = C * vfac * gphi[2];
        const R gg = gphi[0] * gphi[0] + gphi[1] * gphi[1] + gphi[2] * gphi[2];
        if (m.at_w != 0.0 && gg > (R)1e-20) {
            const R jat = -(R)m.at_w * ((R)1 + ((R)1 - (R)m.k_part) * U) * dps / sqrt(gg);
            jx += jat * gphi[0];
            jy += jat * gphi[1];
            jz += jat * gphi[2];
        }
        const R src = (R)m.latent_C * dps;
#pragma unroll
        for (int i = 0; i < 7; i++) {
            const R w = (R)wgt<7>(i);
            const R cj = cx(i) * jx + cy(i) * jy + cz(i) * jz;
            const R ceq = w * (sc + cj / (R)0.25);
            cc[i] = cc[i] - om * (cc[i] - ceq) + w * src;
        }
        FaceState<R> fs;
        fs.dscale = A;  // Dirichlet faces fix U: c_w = A(phi) U_w
        if constexpr (Z)
            zpush_stage<7, R, FACES>(cc, a.c, c, a, kKindScalar, zwc, zmc, tid, act, fs);
        else
            write_pops<7, ODD, R, FACES>(cc, a.c, c, a, kKindScalar, fs);
    } else {
        if (m.need_U) {
            R c7[7];
            read_pops_rt<7, R>(c7, a.c, c, a, kKindScalar);
            R sc = 0;
#pragma unroll
            for (int i = 0; i < 7; i++) sc += c7[i];
            U = ((R)a.c.ref + sc) / A;
        }
    }

    // ---- order parameter j: interaction, source, then its populations
    R P = 0;
    for (int k = 0; k < m.n; k++) {
        if (k != j) {
            const R hk = (R)0.5 * *slot_at(cur + (size_t)k * S, c.id) + (R)0.5;
            P += hk * hk;
        }
    }
    P *= (R)(-2.0 * m.kappa[j]) * ((R)0.5 * pj + (R)0.5);
    R src, zd, n0 = 0, n1 = 0, n2 = 0;
    if constexpr (KR) {
        R Pj[27];
        mpf_stencil<R>(a, c, cur + (size_t)j * S, Pj);
        R Mrot[9];
#pragma unroll
        for (int k = 0; k < 9; k++) Mrot[k] = (R)m.rot[j][k];
        const Aniso<R> an{m.akind[j], (R)m.e4[j], (R)m.ap[j][0], (R)m.ap[j][1]};
        const R upf = m.coupling == 1 ? T : (m.coupling == 2 ? U : (R)0);
        R as;
        const R Sg = phase_source<R>(Pj, upf, an, (R)m.lambda, (R)m.dt, (R)m.wdx2, Mrot,
                                     m.rotated[j] != 0, as);
        src = Sg + (R)m.dt * P / (as * as);
        zd = 0;
    } else {
        R s[3];
        mpf_gradient<R>(a, c, cur + (size_t)j * S, s);
        const R sn = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) + (R)m.eps;
        n0 = s[0] / sn;
        n1 = s[1] / sn;
        n2 = s[2] / sn;
        const R d = (R)1 - pj * pj;
        const R Q = (R)m.growth * (pj - (R)m.lambda * T * d) * d;
        src = (R)m.dt * (P + Q);
        zd = (R)m.zc * d;
    }
    R g[19];
    uint32_t own = 0;
    if constexpr (Z)
        own = zpull_issue<19, R, FACES>(g, L, c, a, zr, tid);
    else
        read_pops<19, ODD, R, FACES>(g, L, c, a, kKindScalar);
    if constexpr (Z) {
        __syncthreads();
        zpull_finish<19, R>(g, c, a, zr, tid, own);
        if constexpr (TH) zpush_finish<7, R, FACES>(a.h, c, a, zwh, zmh, tid);
        if constexpr (SO) zpush_finish<7, R, FACES>(a.c, c, a, zwc, zmc, tid);
    }
    const R ux = (R)m.u[j][0], uy = (R)m.u[j][1], uz = (R)m.u[j][2];
    const R uu = ux * ux + uy * uy + uz * uz;
    const R dp = pj - (R)L.ref;
    const R om = (R)m.omega_g;
#pragma unroll
    for (int i = 0; i < 19; i++) {
        const R w = (R)wgt<19>(i);
        const R cu = cx(i) * ux + cy(i) * uy + cz(i) * uz;
        const R cn = cx(i) * n0 + cy(i) * n1 + cz(i) * n2;
        // deviation form: g~eq = w (phi_j - ref) + w phi_j (3 cu + 4.5 cu^2 - 1.5 uu)
        const R geq = w * (dp + pj * ((R)3 * cu + (R)4.5 * cu * cu - (R)1.5 * uu));
        g[i] = g[i] - om * (g[i] - geq) + w * (src + zd * cn);
    }
    if constexpr (Z) {
        zpush_stage<19, R, FACES>(g, L, c, a, kKindScalar, zw, zm, tid, act, FaceState<R>());
        __syncthreads();
        zpush_finish<19, R, FACES>(L, c, a, zw, zm, tid);
    } else {
        write_pops<19, ODD, R, FACES>(g, L, c, a, kKindScalar);
    }
    };
    if (slow)
        body(std::true_type{});
    else
        body(std::false_type{});
}

// Equilibrium start in the own-slot layout: phase populations from phi_j (doubles, n x N) with
// each order parameter's velocity, both phase buffers, thermal populations from T (or T_ref),
// solute populations from c (or 0).
template <class R>
__global__ void __launch_bounds__(256) k_mpf_init(const MpfArgs m, const double* phij,
                                                  const double* Tin, const double* Cin,
                                                  R* phi0, R* phi1) {
    Cell c;
    if (!locate(m.a, c)) return;
    const uint32_t S = m.a.S, N = m.a.N;
    double comb = m.n - 1.0;
    for (int j = 0; j < m.n; j++) {
        const double pj = phij[(size_t)j * N + c.id];
        R* base = (R*)m.gl[j].pop;
        const double ux = m.u[j][0], uy = m.u[j][1], uz = m.u[j][2];
        const double uu = ux * ux + uy * uy + uz * uz;
#pragma unroll
        for (int i = 0; i < 19; i++) {
            const double w = wgt<19>(i);
            const double cu = cx(i) * ux + cy(i) * uy + cz(i) * uz;
            const double geq = w * pj * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
            *pop_at(base, i, c.id, m.a, m.gl[j]) = (R)(geq - w * m.a.g.ref);
        }
        *slot_at(phi0 + (size_t)j * S, c.id) = (R)pj;
        *slot_at(phi1 + (size_t)j * S, c.id) = (R)pj;
        comb += pj;
    }
    *slot_at(phi0 + (size_t)m.n * S, c.id) = (R)comb;
    *slot_at(phi1 + (size_t)m.n * S, c.id) = (R)comb;
    if (m.thermal) {
        const double t = Tin ? Tin[c.id] : m.T_ref;
        R* hb = (R*)m.a.h.pop;
#pragma unroll
        for (int i = 0; i < 7; i++)
            *pop_at(hb, i, c.id, m.a, m.a.h) = (R)(wgt<7>(i) * (t - m.a.h.ref));
    }
    if (m.a.c.pop) {
        const double cv = Cin ? Cin[c.id] : 0.0;
        R* cb = (R*)m.a.c.pop;
#pragma unroll
        for (int i = 0; i < 7; i++)
            *pop_at(cb, i, c.id, m.a, m.a.c) = (R)(wgt<7>(i) * (cv - m.a.c.ref));
    }
}

template <class R>
int mpf_launch_moments(const MpfArgs& m, int odd, void* phi_out, hipStream_t s) {
    dim3 grid(m.a.nblk), block(256);
    const bool per = !m.a.has_walls && m.a.g.npmask == 0;
    if (odd && m.a.zstage) {
        if (per) k_mpf_moments<R, true, true, true><<<grid, block, 0, s>>>(m, (R*)phi_out);
        else k_mpf_moments<R, true, true><<<grid, block, 0, s>>>(m, (R*)phi_out);
    } else if (odd) {
        k_mpf_moments<R, true><<<grid, block, 0, s>>>(m, (R*)phi_out);
    } else {
        k_mpf_moments<R, false><<<grid, block, 0, s>>>(m, (R*)phi_out);
    }
    return 0;
}

}  // namespace mkg

using namespace mkg;

struct mkg_mpf_ctx {
    mkg_mpf_params p;
    int fp32 = 0;
    size_t esz = 8;
    uint32_t N = 0, S = 0, F = 0, nblk = 0;
    hipStream_t stream = nullptr;
    // population block: the n order-parameter lattices (block lattices 0..n-1), the thermal
    // (n) and the solute (n + 1) lattice; h / c point at the thermal / solute slot 0 (null when
    // absent)
    void* pops = nullptr;
    PopBlock pb;
    void* h = nullptr;
    void* c = nullptr;
    void* phi[2] = {nullptr, nullptr};  // (n + 1) S elements each
    int8_t* flag = nullptr;
    double* scratch = nullptr;  // 19 N doubles
    double* partials = nullptr;
    double* red_out = nullptr;
    int lay = 0;              // AA layout of every lattice
    int zstage = 0;           // O steps take the z shift through LDS
    int xcd_chunk = 0;        // block order of the launches, fixed at creation
    FaceTab* ftab = nullptr;  // boundary table of the face rules
    int64_t pf_count = 0;
    int has_walls = 0;
    // order parameter j's lattice descriptor (base, face buffer) on top of a filled one
    void lattice(LatArgs& L, int j) const { pb.lattice(L, pops, j, esz); }
};

namespace {

int mfail(const std::string& msg) {
    set_error(msg);
    return -1;
}

#define MCHK(expr)                                                                     \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return mfail(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr uint32_t kMpfPad = 1024;

MpfArgs mpf_args(const mkg_mpf_ctx* h) {
    MpfArgs m;
    std::memset(&m, 0, sizeof(m));
    const mkg_mpf_params& p = h->p;
    StepArgs& a = m.a;
    a.nx = p.nx;
    a.ny = p.ny;
    a.nz = p.nz;
    a.N = h->N;
    a.S = h->S;
    a.nblk = h->nblk;
    a.F = h->F;
    a.has_walls = h->has_walls;
    a.xcd_chunk = h->xcd_chunk;
    a.zstage = h->zstage;
    a.flag = h->flag;
    a.rho0 = 1.0;
    h->pb.apply(a);
    auto fill = [&](LatArgs& L, int k, const int32_t* bc, double ref, const double* fval) {
        h->pb.lattice(L, h->pops, k, h->esz);
        int np = 0;
        for (int f = 0; f < 6; f++) {
            L.bc[f] = bc[f];
            if (bc[f] != MKG_BC_PERIODIC) np |= 1 << f;
            L.fval[f] = fval[f];
        }
        L.npmask = np;
        L.lay = h->lay;
        L.adv = 1;
        L.ref = ref;
    };
    fill(a.g, 0, p.bc, -1.0, p.phi_face);
    a.g.fid = 1;  // every order parameter has the phase lattice's faces
    for (int j = 0; j < MKG_MPF_MAX; j++) {
        m.gl[j] = a.g;
        if (j < p.n) h->lattice(m.gl[j], j);
    }
    if (h->h) fill(a.h, p.n, p.bc_h, p.T_ref, p.T_face);
    if (h->c) fill(a.c, p.n + 1, p.bc_c, 0.0, p.C_face);
    a.h.fid = 2;
    a.c.fid = 3;
    a.ftab = h->ftab;
    a.phi_cur = h->phi[h->pf_count & 1];
    a.phi_prev = h->phi[(h->pf_count + 1) & 1];
    m.n = p.n;
    m.thermal = p.thermal ? 1 : 0;
    for (int j = 0; j < MKG_MPF_MAX; j++) {
        for (int k = 0; k < 3; k++) m.u[j][k] = p.u[j][k];
        m.kappa[j] = p.kappa[j];
    }
    m.growth = p.growth;
    m.lambda = p.lambda;
    m.eps = p.eps;
    m.omega_g = 1.0 / (0.5 + p.dt * p.wdx * p.wdx * 3.0);
    m.zc = (1.0 - 0.5 * m.omega_g) / (std::sqrt(2.0) * p.wdx);
    m.dt = p.dt;
    m.omega_h = 1.0 / (0.5 + p.alpha / 0.25);
    m.latent_T = p.latent_T;
    m.T_ref = p.T_ref;
    const bool kr = p.model == MKG_MPF_KARMA_RAPPEL;
    m.coupling = kr ? p.coupling : 0;
    m.need_T = h->h && (kr ? p.coupling == 1 : p.growth != 0.0);
    m.need_U = h->c && kr && p.coupling == 2;
    for (int j = 0; j < MKG_MPF_MAX; j++) {
        m.akind[j] = p.aniso_kind[j];
        m.e4[j] = p.eps4[j];
        m.ap[j][0] = p.aniso[j][0];
        m.ap[j][1] = p.aniso[j][1];
        int ident = 1;
        for (int k = 0; k < 9; k++) {
            m.rot[j][k] = p.rot[j][k];
            ident &= p.rot[j][k] == ((k % 4 == 0) ? 1.0 : 0.0);
        }
        m.rotated[j] = !ident;
    }
    m.wdx2 = p.wdx * p.wdx;
    m.D_l = p.D_l;
    m.D_s = p.D_s;
    m.k_part = h->c ? p.k_part : 1.0;
    m.q_min = p.q_min;
    m.latent_C = p.latent_C;
    m.at_w = p.at_coef * p.wdx;
    return m;
}

// Rewrite every lattice into the own-slot layout.
int mpf_normalize(mkg_mpf_ctx* h) {
    if (h->lay == 0) return 0;
    MpfArgs m = mpf_args(h);
    for (int j = 0; j < h->p.n; j++) {
        LatArgs L = m.a.g;
        h->lattice(L, j);
        if (launch_canonical(19, h->fp32, m.a, L, 1, h->scratch, h->stream) ||
            launch_store_pops(19, h->fp32, m.a, L, h->scratch, h->stream))
            return mfail("normalize: lattice unavailable");
    }
    for (const LatArgs* S7 : {&m.a.h, &m.a.c}) {
        if (!S7->pop) continue;
        if (launch_canonical(7, h->fp32, m.a, *S7, 1, h->scratch, h->stream) ||
            launch_store_pops(7, h->fp32, m.a, *S7, h->scratch, h->stream))
            return mfail("normalize: lattice unavailable");
    }
    MCHK(hipGetLastError());
    h->lay = 0;
    return 0;
}

template <class R, bool TH, bool SO, bool KR>
void launch_mpf_one(const MpfArgs& m, int j, int odd, int zs, hipStream_t s) {
    dim3 grid(m.a.nblk), block(256);
    if (odd && zs)
        k_mpf_one<R, TH, SO, KR, true, true><<<grid, block, 0, s>>>(m, m.gl[j], j);
    else if (odd)
        k_mpf_one<R, TH, SO, KR, true><<<grid, block, 0, s>>>(m, m.gl[j], j);
    else
        k_mpf_one<R, TH, SO, KR, false><<<grid, block, 0, s>>>(m, m.gl[j], j);
}
// Collide-stream of every lattice: one launch per order parameter, the scalar lattices on the
// last.  The solute lattice exists only with Karma-Rappel grains.
template <class R>
void launch_mpf_step_kernel(const MpfArgs& m, int odd, int th, int so, int kr, hipStream_t s) {
    const int zs = m.a.zstage;
    for (int j = 0; j < m.n; j++) {
        const bool last = j == m.n - 1;
        if (!kr) {
            if (th && last) launch_mpf_one<R, true, false, false>(m, j, odd, zs, s);
            else launch_mpf_one<R, false, false, false>(m, j, odd, zs, s);
        } else if (last && th && so) {
            launch_mpf_one<R, true, true, true>(m, j, odd, zs, s);
        } else if (last && th) {
            launch_mpf_one<R, true, false, true>(m, j, odd, zs, s);
        } else if (last && so) {
            launch_mpf_one<R, false, true, true>(m, j, odd, zs, s);
        } else {
            launch_mpf_one<R, false, false, true>(m, j, odd, zs, s);
        }
    }
}

void mpf_destroy_impl(mkg_mpf_ctx* h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void* bufs[] = {h->pops, h->phi[0], h->phi[1], h->flag, h->scratch, h->partials,
                    h->red_out, h->ftab};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

}  // namespace

extern "C" {

size_t mkg_mpf_params_size(void) { return sizeof(mkg_mpf_params); }

int mkg_mpf_params_default(mkg_mpf_params* m) {
    if (!m) return mfail("params is null");
    std::memset(m, 0, sizeof(*m));
    m->nx = m->ny = m->nz = 1;
    m->n = 2;
    m->dt = 0.008;
    m->wdx = 2.5;
    m->u[1][0] = 0.02;
    for (int j = 1; j < MKG_MPF_MAX; j++) m->kappa[j] = 1.0;
    m->eps = 1e-12;
    m->alpha = 0.1;
    m->latent_T = 0.5;
    for (int f = 0; f < 6; f++) m->phi_face[f] = -1.0;
    for (int j = 0; j < MKG_MPF_MAX; j++)
        for (int k = 0; k < 3; k++) m->rot[j][4 * k] = 1.0;
    m->k_part = 1.0;
    m->q_min = 1e-3;
    m->latent_C = 0.5;
    return 0;
}

int mkg_mpf_create(const mkg_mpf_params* p, mkg_mpf_handle* out) {
    if (!out) return mfail("out is null");
    *out = nullptr;
    if (!p) return mfail("params is null");
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) return mfail("grid dimensions must be >= 1");
    if (p->n < 1 || p->n > MKG_MPF_MAX) return mfail("n must be 1..MKG_MPF_MAX");
    if (p->precision != MKG_PREC_FP64 && p->precision != MKG_PREC_FP32)
        return mfail("precision must be MKG_PREC_FP64 or MKG_PREC_FP32");
    if (!(p->wdx > 0) || !(p->dt > 0)) return mfail("wdx and dt must be > 0");
    if (!(p->eps > 0)) return mfail("eps must be > 0");
    const double esz = p->precision == MKG_PREC_FP32 ? 4.0 : 8.0;
    const double n = (double)p->nx * p->ny * p->nz;
    if ((n + 65536.0) * esz * (p->n + 1) >= 4294967296.0)
        return mfail("grid too large: the phase buffer (cells x (n+1) x element size) must stay below 4 GiB");
    for (int f = 0; f < 6; f++) {
        const int32_t codes[3] = {p->bc[f], p->bc_h[f], p->bc_c[f]};
        for (int32_t v : codes)
            if (v < MKG_BC_PERIODIC || v > MKG_BC_OUTFLOW)
                return mfail("face codes of the multi-phase field are periodic, bounce, Dirichlet or outflow");
    }
    for (int a = 0; a < 3; a++)
        if ((p->bc[2 * a] == 0) != (p->bc[2 * a + 1] == 0) ||
            (p->bc_h[2 * a] == 0) != (p->bc_h[2 * a + 1] == 0) ||
            (p->bc_c[2 * a] == 0) != (p->bc_c[2 * a + 1] == 0))
            return mfail("a periodic axis needs periodic codes on both faces");
    if (p->model != MKG_MPF_TEMPLATE && p->model != MKG_MPF_KARMA_RAPPEL)
        return mfail("model must be MKG_MPF_TEMPLATE or MKG_MPF_KARMA_RAPPEL");
    if (p->solute && p->model != MKG_MPF_KARMA_RAPPEL)
        return mfail("the solute lattice drives Karma-Rappel grains (model MKG_MPF_KARMA_RAPPEL)");
    if (p->model == MKG_MPF_KARMA_RAPPEL) {
        if (p->coupling < 0 || p->coupling > 2) return mfail("coupling must be 0, 1 (T) or 2 (U)");
        if (p->coupling == 1 && !p->thermal) return mfail("coupling 1 needs the thermal lattice");
        if (p->coupling == 2 && !p->solute) return mfail("coupling 2 needs the solute lattice");
        for (int j = 0; j < p->n; j++) {
            if (p->aniso_kind[j] < MKG_ANISO_CUBIC || p->aniso_kind[j] > MKG_ANISO_FACET)
                return mfail("aniso_kind of every grain must be one of MKG_ANISO_*");
            double err = 0;  // rotations: R^T R = I
            for (int r = 0; r < 3; r++)
                for (int q = 0; q < 3; q++) {
                    double d = 0;
                    for (int k = 0; k < 3; k++) d += p->rot[j][3 * k + r] * p->rot[j][3 * k + q];
                    err = std::fmax(err, std::fabs(d - (r == q ? 1.0 : 0.0)));
                }
            if (!(err < 1e-9)) return mfail("rot of every grain must be an orthogonal matrix");
        }
    }
    if (p->solute) {
        if (!(p->D_l >= 0) || !(p->D_s >= 0)) return mfail("D_l and D_s must be >= 0");
        if (!(p->k_part > 0) || !(p->k_part <= 1)) return mfail("k_part must lie in (0, 1]");
        if (!(p->q_min > 0)) return mfail("q_min must be > 0");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return mfail("no HIP device available");
    mkg_mpf_ctx* h = new mkg_mpf_ctx();
    h->p = *p;
    h->fp32 = p->precision == MKG_PREC_FP32;
    h->esz = h->fp32 ? 4 : 8;
    h->N = (uint32_t)n;
    h->S = ((h->N + 127u) & ~127u) + kMpfPad;
    h->F = 2u * ((uint32_t)p->ny * p->nz + (uint32_t)p->nx * p->nz + (uint32_t)p->nx * p->ny);
    h->nblk = (h->N + 255u) / 256u;
    // z stage for O steps of the collide-stream when a 256-thread block holds whole z rows
    // (nz = 64, 128 or 256); MKG_ZSTAGE=0 turns it off
    h->zstage = (p->nz >= 64 && 256 % p->nz == 0) ? 1 : 0;
    if (const char* env = std::getenv("MKG_ZSTAGE")) h->zstage = h->zstage && std::atoi(env) != 0;
    h->xcd_chunk = xcd_chunk_setting(h->esz);
    auto bad = [&](hipError_t e, const char* what) {
        mpf_destroy_impl(h);
        return mfail(std::string(what) + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipStreamCreate(&h->stream)) != hipSuccess) return bad(e, "hipStreamCreate");
    {
        int q[kMaxLattices] = {};
        for (int j = 0; j < p->n; j++) q[j] = 19;
        q[p->n] = p->thermal ? 7 : 0;
        q[p->n + 1] = p->solute ? 7 : 0;
        h->pb = plan_pop_block(q, p->n + 2, h->N, h->S, h->F, h->esz);
        const size_t gb = h->pb.total * h->esz;
        if ((e = hipMalloc(&h->pops, gb)) != hipSuccess) return bad(e, "hipMalloc");
        if ((e = hipMemsetAsync(h->pops, 0, gb, h->stream)) != hipSuccess) return bad(e, "hipMemsetAsync");
        if (p->thermal) h->h = (char*)h->pops + h->pb.slot0[p->n] * h->esz;
        if (p->solute) h->c = (char*)h->pops + h->pb.slot0[p->n + 1] * h->esz;
    }
    const size_t pb = (size_t)(p->n + 1) * h->S * h->esz;
    for (int b = 0; b < 2; b++) {
        if ((e = hipMalloc(&h->phi[b], pb)) != hipSuccess) return bad(e, "hipMalloc");
        if ((e = hipMemsetAsync(h->phi[b], 0, pb, h->stream)) != hipSuccess) return bad(e, "hipMemsetAsync");
    }
    if ((e = hipMalloc((void**)&h->flag, h->S)) != hipSuccess) return bad(e, "hipMalloc");
    if ((e = hipMemsetAsync(h->flag, 0, h->S, h->stream)) != hipSuccess) return bad(e, "hipMemsetAsync");
    if ((e = hipMalloc((void**)&h->scratch, (size_t)19 * h->N * sizeof(double))) != hipSuccess)
        return bad(e, "hipMalloc");
    if ((e = hipMalloc((void**)&h->partials, (size_t)(h->nblk + 1) * 4 * sizeof(double))) != hipSuccess)
        return bad(e, "hipMalloc");
    if ((e = hipMalloc((void**)&h->red_out, 8 * sizeof(double))) != hipSuccess) return bad(e, "hipMalloc");
    if ((e = facetab_alloc(&h->ftab)) != hipSuccess) return bad(e, "hipMalloc");
    if ((e = facetab_upload(h->ftab, mpf_args(h).a, h->stream)) != hipSuccess)
        return bad(e, "boundary table upload");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return bad(e, "hipStreamSynchronize");
    *out = h;
    return 0;
}

int mkg_mpf_destroy(mkg_mpf_handle h) {
    mpf_destroy_impl(h);
    return 0;
}

int mkg_mpf_init(mkg_mpf_handle h, const double* phi, const double* T) {
    return mkg_mpf_init_fields(h, phi, T, nullptr);
}

int mkg_mpf_init_fields(mkg_mpf_handle h, const double* phi, const double* T, const double* C) {
    if (!h) return mfail("null handle");
    if (!phi) return mfail("phi is null");
    const size_t N = h->N;
    double* dphi = nullptr;
    double* dT = nullptr;
    double* dC = nullptr;
    MCHK(hipMalloc((void**)&dphi, (size_t)h->p.n * N * sizeof(double)));
    hipError_t e = hipMemcpyAsync(dphi, phi, (size_t)h->p.n * N * sizeof(double),
                                  hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess && T && h->h) {
        e = hipMalloc((void**)&dT, N * sizeof(double));
        if (e == hipSuccess)
            e = hipMemcpyAsync(dT, T, N * sizeof(double), hipMemcpyHostToDevice, h->stream);
    }
    if (e == hipSuccess && C && h->c) {
        e = hipMalloc((void**)&dC, N * sizeof(double));
        if (e == hipSuccess)
            e = hipMemcpyAsync(dC, C, N * sizeof(double), hipMemcpyHostToDevice, h->stream);
    }
    if (e == hipSuccess) {
        h->lay = 0;
        h->pf_count = 0;
        MpfArgs m = mpf_args(h);
        dim3 grid(h->nblk), block(256);
        if (h->fp32)
            k_mpf_init<float><<<grid, block, 0, h->stream>>>(m, dphi, dT, dC, (float*)h->phi[0],
                                                             (float*)h->phi[1]);
        else
            k_mpf_init<double><<<grid, block, 0, h->stream>>>(m, dphi, dT, dC,
                                                              (double*)h->phi[0],
                                                              (double*)h->phi[1]);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dphi);
    if (dT) (void)hipFree(dT);
    if (dC) (void)hipFree(dC);
    if (e != hipSuccess) return mfail(std::string("mkg_mpf_init: ") + hipGetErrorString(e));
    return 0;
}

int mkg_mpf_step(mkg_mpf_handle h, int nsteps) {
    if (!h) return mfail("null handle");
    if (nsteps < 0) return mfail("nsteps must be >= 0");
    for (int s = 0; s < nsteps; s++) {
        h->pf_count++;
        MpfArgs m = mpf_args(h);
        const int odd = h->lay;
        if (h->fp32) mpf_launch_moments<float>(m, odd, h->phi[h->pf_count & 1], h->stream);
        else mpf_launch_moments<double>(m, odd, h->phi[h->pf_count & 1], h->stream);
        const int th = h->h != nullptr, so = h->c != nullptr;
        const int kr = h->p.model == MKG_MPF_KARMA_RAPPEL;
        if (h->fp32) launch_mpf_step_kernel<float>(m, odd, th, so, kr, h->stream);
        else launch_mpf_step_kernel<double>(m, odd, th, so, kr, h->stream);
        h->lay ^= 1;
    }
    MCHK(hipGetLastError());
    return 0;
}

int mkg_mpf_sync(mkg_mpf_handle h) {
    if (!h) return mfail("null handle");
    MCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int mkg_mpf_get(mkg_mpf_handle h, int which, void* host, size_t count) {
    if (!h) return mfail("null handle");
    if (!host) return mfail("host buffer is null");
    const size_t N = h->N;
    const int n = h->p.n;
    MpfArgs m = mpf_args(h);
    if (which == MKG_MPF_POPS || which == MKG_MPF_THERMAL || which == MKG_MPF_SOLUTE) {
        const int nl = which == MKG_MPF_POPS ? n : 1;
        const int q = which == MKG_MPF_POPS ? 19 : 7;
        if (which == MKG_MPF_THERMAL && !h->h) return mfail("thermal lattice not present");
        if (which == MKG_MPF_SOLUTE && !h->c) return mfail("solute lattice not present");
        if (count != (size_t)nl * q * N) return mfail("count does not match the field size");
        for (int j = 0; j < nl; j++) {
            LatArgs L = which == MKG_MPF_POPS ? m.a.g : (which == MKG_MPF_THERMAL ? m.a.h : m.a.c);
            if (which == MKG_MPF_POPS) h->lattice(L, j);
            if (launch_canonical(q, h->fp32, m.a, L, 1, h->scratch, h->stream))
                return mfail("lattice unavailable");
            MCHK(hipMemcpyAsync((double*)host + (size_t)j * q * N, h->scratch,
                                (size_t)q * N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        }
        MCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    if (which == MKG_MPF_PHI || which == MKG_MPF_PHI_PREV) {
        if (count != (size_t)(n + 1) * N) return mfail("count does not match the field size");
        const void* src = which == MKG_MPF_PHI ? m.a.phi_cur : m.a.phi_prev;
        for (int j = 0; j <= n; j++) {
            launch_to_double(h->fp32, (const char*)src + (size_t)j * h->S * h->esz, h->scratch,
                             h->N, h->stream);
            MCHK(hipMemcpyAsync((double*)host + (size_t)j * N, h->scratch, N * sizeof(double),
                                hipMemcpyDeviceToHost, h->stream));
        }
        MCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    if (which == MKG_MPF_FLAG) {
        if (count != N) return mfail("count does not match the field size");
        MCHK(hipMemcpyAsync(host, h->flag, N, hipMemcpyDeviceToHost, h->stream));
        MCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    return mfail("unknown multi-phase field id");
}

int mkg_mpf_set(mkg_mpf_handle h, int which, const void* host, size_t count) {
    if (!h) return mfail("null handle");
    if (!host) return mfail("host buffer is null");
    const size_t N = h->N;
    const int n = h->p.n;
    if (which == MKG_MPF_POPS || which == MKG_MPF_THERMAL || which == MKG_MPF_SOLUTE) {
        const int nl = which == MKG_MPF_POPS ? n : 1;
        const int q = which == MKG_MPF_POPS ? 19 : 7;
        if (which == MKG_MPF_THERMAL && !h->h) return mfail("thermal lattice not present");
        if (which == MKG_MPF_SOLUTE && !h->c) return mfail("solute lattice not present");
        if (count != (size_t)nl * q * N) return mfail("count does not match the field size");
        if (mpf_normalize(h)) return -1;
        MpfArgs m = mpf_args(h);
        for (int j = 0; j < nl; j++) {
            LatArgs L = which == MKG_MPF_POPS ? m.a.g : (which == MKG_MPF_THERMAL ? m.a.h : m.a.c);
            if (which == MKG_MPF_POPS) h->lattice(L, j);
            MCHK(hipMemcpyAsync(h->scratch, (const double*)host + (size_t)j * q * N,
                                (size_t)q * N * sizeof(double), hipMemcpyHostToDevice, h->stream));
            if (launch_store_pops(q, h->fp32, m.a, L, h->scratch, h->stream))
                return mfail("lattice unavailable");
        }
        MCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    if (which == MKG_MPF_PHI || which == MKG_MPF_PHI_PREV) {
        if (count != (size_t)(n + 1) * N) return mfail("count does not match the field size");
        MpfArgs m = mpf_args(h);
        void* dst = (void*)(which == MKG_MPF_PHI ? m.a.phi_cur : m.a.phi_prev);
        for (int j = 0; j <= n; j++) {
            MCHK(hipMemcpyAsync(h->scratch, (const double*)host + (size_t)j * N,
                                N * sizeof(double), hipMemcpyHostToDevice, h->stream));
            launch_from_double(h->fp32, h->scratch, (char*)dst + (size_t)j * h->S * h->esz,
                               h->N, h->stream);
        }
        MCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    if (which == MKG_MPF_FLAG) {
        if (count != N) return mfail("count does not match the field size");
        const int8_t* f = (const int8_t*)host;
        int walls = 0;
        for (size_t k = 0; k < N; k++) {
            if (f[k] != kIsFluid && f[k] != kIsWalls)
                return mfail("multi-phase flags are 0 (fluid) or 3 (wall)");
            walls |= f[k] == kIsWalls;
        }
        if (mpf_normalize(h)) return -1;
        MCHK(hipMemcpyAsync(h->flag, host, N, hipMemcpyHostToDevice, h->stream));
        MCHK(hipStreamSynchronize(h->stream));
        h->has_walls = walls;
        return 0;
    }
    return mfail("unknown multi-phase field id");
}

int mkg_mpf_sums(mkg_mpf_handle h, double* out) {
    if (!h) return mfail("null handle");
    if (!out) return mfail("out is null");
    MpfArgs m = mpf_args(h);
    for (int j = 0; j <= h->p.n; j++) {
        launch_to_double(h->fp32, (const char*)m.a.phi_cur + (size_t)j * h->S * h->esz,
                         h->scratch, h->N, h->stream);
        int nout = 0;
        launch_reduce(4, m.a, nullptr, nullptr, nullptr, nullptr, h->scratch, h->partials,
                      h->red_out, &nout, h->stream);
        MCHK(hipMemcpyAsync(out + j, h->red_out, sizeof(double), hipMemcpyDeviceToHost,
                            h->stream));
    }
    MCHK(hipStreamSynchronize(h->stream));
    return 0;
}

}  // extern "C"
