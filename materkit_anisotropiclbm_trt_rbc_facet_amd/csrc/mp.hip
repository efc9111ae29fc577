// Phase-field two-phase flow of mkFluidsMultiphase (MaterKit/inc/mkFluidsMultiphase.h:20-99;
// run_flow_multiphase, MaterKit/MaterKit.cpp:290-331).  Its parameters (rho_l, rho_g, sigma,
// W_0, beta = 12 sigma/W, kappa = 1.5 sigma W, eta, mkFluidsMultiphase.h:81-97) are those of the
// model of Liang et al. (PRE 89, 053320, 2014), restated here with BGK collisions: a
// Cahn-Hilliard lattice h for the order parameter phi in [0, 1] and a velocity-pressure flow
// lattice f, both D3Q19, in the AA storage and face rules of the fused solver (device.h).
//
// One step is three launches:
//   k_mp_phi   phi = sum_i h_i
//   k_mp_mu    mu = 4 beta phi (phi - 1)(phi - 1/2) - kappa lap phi (isotropic 19-point Laplacian)
//   k_mp_step  per cell: grad phi; rho = rho_g + phi (rho_l - rho_g); F = mu grad phi + rho g +
//              force; rho u = sum c f + F/2; p = cs^2/(1 - w0) [sum_{i>0} f + u.grad rho/2 +
//              rho s_0(u)]; flow BGK with tau_f = nu(phi)/cs^2 + 1/2 and the forcing
//              (1 - 1/(2 tau_f)) w_i [c_i.F + (c_i.u)(c_i.grad rho)]/cs^2 (its zeroth moment
//              u.grad rho keeps div u = 0 where rho varies); nu(phi) per visc_model;
//              Cahn-Hilliard BGK with h0eq = phi + (w0 - 1) eta mu, hieq = w_i (eta mu +
//              phi c_i.u/cs^2) and the source (1 - 1/(2 tau_h)) w_i c_i.d(phi u)/cs^2, where
//              d(phi u) is the change of phi u over the step (kept per cell)
// Wetting walls (initialize_wetting, mkFluidsMultiphase.h:35; theta, :95): the surface energy
// -sigma cos(theta) phi^2 (3 - 2 phi) gives n.grad phi = (4/W) cos(theta) phi (1 - phi) along
// the outward normal, imposed through the ghost value phi_e + (4/W) cos(theta) phi_e (1 - phi_e)
// one cell across the wall in every phi stencil (mu, grad phi, grad rho).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/mkg.h"
#include "device.h"

namespace mkg {

void set_error(const std::string& msg);

struct MpArgs {
    StepArgs a;  // geometry and faces; a.f = flow lattice, a.g = Cahn-Hilliard lattice
    double rho_l, rho_g, nu_l, nu_g, beta, kappa, eta, tau_h;
    double gravity[3], force[3];
    void* phi;   // N (storage type)
    void* mu;    // N
    void* phiu;  // 3 planes of S: phi u of the last step
    double wet_k;      // (4/W) cos(theta)
    uint32_t wet_mask;  // face bits (XLO..ZHI) of the wetting walls
    int32_t visc;       // MKG_MP_VISC_*
};

// Kinematic viscosity at a clamped phi (MKG_MP_VISC_* of mkg.h).
template <class R>
__device__ inline R mp_nu(const MpArgs& m, R pc) {
    if (m.visc == MKG_MP_VISC_NU_LINEAR) return (R)m.nu_g + pc * (R)(m.nu_l - m.nu_g);
    const R rho = (R)m.rho_g + pc * (R)(m.rho_l - m.rho_g);
    const R ml = (R)(m.rho_l * m.nu_l), mg = (R)(m.rho_g * m.nu_g);
    const R mu = m.visc == MKG_MP_VISC_MU_LINEAR ? mg + pc * (ml - mg)
                                                 : (R)1 / (pc / ml + ((R)1 - pc) / mg);
    return mu / rho;
}

// 19-point stencil of phi (periodic faces wrap, others replicate the edge; across a wetting
// wall the edge value phi_e becomes phi_e + wet_k phi_e (1 - phi_e), once per such face).
template <class R>
__device__ inline void mp_stencil(const MpArgs& m, const Cell& c, const R* v, R (&P)[19]) {
    const StepArgs& a = m.a;
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
    for (int i = 0; i < 19; i++) P[i] = *slot_at(v, px[cx(i) + 1] + py[cy(i) + 1] + pz[cz(i) + 1]);
    const uint32_t wm = m.wet_mask & c.bits;
    if (wm) {
        const R k = (R)m.wet_k;
#pragma unroll
        for (int i = 1; i < 19; i++) {
            const int n = ((cx(i) < 0) & !!(wm & XLO)) + ((cx(i) > 0) & !!(wm & XHI)) +
                          ((cy(i) < 0) & !!(wm & YLO)) + ((cy(i) > 0) & !!(wm & YHI)) +
                          ((cz(i) < 0) & !!(wm & ZLO)) + ((cz(i) > 0) & !!(wm & ZHI));
            if (n) P[i] += (R)n * k * P[i] * ((R)1 - P[i]);
        }
    }
}

// ZS (with ODD): the block holds whole z rows, and the z-moving populations are loaded at the
// cell's own z and shifted through LDS (every thread takes part in the barrier).  PER: no face
// rules (every face of the lattice periodic, no wall cells).
template <class R, bool ODD, bool ZS = false, bool PER = false>
__global__ void __launch_bounds__(256) k_mp_phi(const MpArgs m) {
    constexpr bool Z = ODD && ZS;
    __shared__ R zv[Z ? nzdirs<19>() : 1][256];
    Cell c;
    R h[19];
    if constexpr (Z) {
        const bool act = locate_clamped(m.a, c);
        const uint32_t own = zpull_issue<19, R, !PER>(h, m.a.g, c, m.a, zv, threadIdx.x);
        __syncthreads();
        zpull_finish<19, R>(h, c, m.a, zv, threadIdx.x, own);
        if (!act) return;
    } else {
        if (!locate(m.a, c)) return;
        read_pops<19, ODD, R>(h, m.a.g, c, m.a, kKindScalar);
    }
    R s = 0;
#pragma unroll
    for (int i = 0; i < 19; i++) s += h[i];
    *slot_at((R*)m.phi, c.id) = s;
}

template <class R>
__global__ void __launch_bounds__(256) k_mp_mu(const MpArgs m) {
    Cell c;
    if (!locate(m.a, c)) return;
    R P[19];
    mp_stencil<R>(m, c, (const R*)m.phi, P);
    R lap = 0;
#pragma unroll
    for (int i = 1; i < 19; i++) lap += (R)(6.0 * wgt<19>(i)) * (P[i] - P[0]);
    const R ph = P[0];
    *slot_at((R*)m.mu, c.id) =
        (R)(4.0 * m.beta) * ph * (ph - (R)1) * (ph - (R)0.5) - (R)m.kappa * lap;
}

// k_mp_mu when a block holds whole z rows (m.a.zstage): the nine (dx, dy) columns of the
// stencil are read at the cell's own z (aligned loads), and the z -+ 1 values of the five
// columns that need them come from the neighbouring threads of the row through LDS: wrapped on
// a periodic z face, the cell's own z on a non-periodic one (the edge replication of
// mp_stencil).  Nine loads per cell instead of nineteen, ten of them misaligned.  Every thread
// of the block takes part (one barrier).
template <class R>
__global__ void __launch_bounds__(256) k_mp_mu_z(const MpArgs m) {
    __shared__ R zs[5][256];
    const StepArgs& a = m.a;
    Cell c;
    const bool valid = locate_clamped(a, c);
    const uint32_t tid = threadIdx.x;
    const R* __restrict__ v = (const R*)m.phi;
    const uint32_t np = (uint32_t)a.g.npmask & c.bits;
    uint32_t px[3] = {c.ox[0], c.ox[1], c.ox[2]};
    uint32_t py[3] = {c.oy[0], c.oy[1], c.oy[2]};
    if (np & XLO) px[0] = px[1];
    if (np & XHI) px[2] = px[1];
    if (np & YLO) py[0] = py[1];
    if (np & YHI) py[2] = py[1];
    R col[3][3];
#pragma unroll
    for (int dx = 0; dx < 3; dx++)
#pragma unroll
        for (int dy = 0; dy < 3; dy++) col[dx][dy] = *slot_at(v, px[dx] + py[dy] + c.oz[1]);
    // the columns with z neighbours in the stencil: (0,0), (-1,0), (1,0), (0,-1), (0,1)
    constexpr int CX[5] = {0, -1, 1, 0, 0}, CY[5] = {0, 0, 0, -1, 1};
#pragma unroll
    for (int k = 0; k < 5; k++) zs[k][tid] = col[CX[k] + 1][CY[k] + 1];
    __syncthreads();
    if (!valid) return;
    const uint32_t zm1 = (uint32_t)a.nz - 1u;
    const uint32_t rb = tid - (tid & zm1);
    const uint32_t z = (uint32_t)c.z;
    const uint32_t zlo = rb + ((np & ZLO) ? z : ((z - 1u) & zm1));
    const uint32_t zhi = rb + ((np & ZHI) ? z : ((z + 1u) & zm1));
    R P[19];
#pragma unroll
    for (int i = 0; i < 19; i++) {
        if (cz(i) == 0) {
            P[i] = col[cx(i) + 1][cy(i) + 1];
        } else {
#pragma unroll
            for (int k = 0; k < 5; k++)
                if (CX[k] == cx(i) && CY[k] == cy(i)) P[i] = zs[k][cz(i) < 0 ? zlo : zhi];
        }
    }
    const uint32_t wm = m.wet_mask & c.bits;
    if (wm) {
        const R kw = (R)m.wet_k;
#pragma unroll
        for (int i = 1; i < 19; i++) {
            const int n = ((cx(i) < 0) & !!(wm & XLO)) + ((cx(i) > 0) & !!(wm & XHI)) +
                          ((cy(i) < 0) & !!(wm & YLO)) + ((cy(i) > 0) & !!(wm & YHI)) +
                          ((cz(i) < 0) & !!(wm & ZLO)) + ((cz(i) > 0) & !!(wm & ZHI));
            if (n) P[i] += (R)n * kw * P[i] * ((R)1 - P[i]);
        }
    }
    R lap = 0;
#pragma unroll
    for (int i = 1; i < 19; i++) lap += (R)(6.0 * wgt<19>(i)) * (P[i] - P[0]);
    const R ph = P[0];
    *slot_at((R*)m.mu, c.id) =
        (R)(4.0 * m.beta) * ph * (ph - (R)1) * (ph - (R)0.5) - (R)m.kappa * lap;
}

// Macroscopic state of a cell from its pre-collision f: rho, u, p, F and grad rho.
template <class R>
__device__ inline void mp_macro(const MpArgs& m, const R (&f)[19], R ph, R muc, const R (&g)[3],
                                R& rho, R (&u)[3], R& p, R (&F)[3], R (&gr)[3]) {
    rho = (R)m.rho_g + ph * (R)(m.rho_l - m.rho_g);
    R j[3] = {0, 0, 0}, s1 = 0;
#pragma unroll
    for (int i = 1; i < 19; i++) {
        s1 += f[i];
        if (cx(i)) j[0] += (R)cx(i) * f[i];
        if (cy(i)) j[1] += (R)cy(i) * f[i];
        if (cz(i)) j[2] += (R)cz(i) * f[i];
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        gr[d] = (R)(m.rho_l - m.rho_g) * g[d];
        F[d] = muc * g[d] + rho * (R)m.gravity[d] + (R)m.force[d];
        u[d] = (j[d] + (R)0.5 * F[d]) / rho;
    }
    const R uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const R s0 = (R)wgt<19>(0) * ((R)-1.5 * uu);
    const R ugr = u[0] * gr[0] + u[1] * gr[1] + u[2] * gr[2];
    p = (R)(1.0 / 3.0 / (1.0 - wgt<19>(0))) * (s1 + (R)0.5 * ugr + rho * s0);
}

template <class R>
__device__ inline void mp_gradient(const R (&P)[19], R (&g)[3]) {
    g[0] = g[1] = g[2] = 0;
#pragma unroll
    for (int i = 1; i < 19; i++) {
        const R v = (R)(3.0 * wgt<19>(i)) * P[i];
        if (cx(i)) g[0] += (R)cx(i) * v;
        if (cy(i)) g[1] += (R)cy(i) * v;
        if (cz(i)) g[2] += (R)cz(i) * v;
    }
}

// ZS: O step with the z stage of the fused step (the block holds whole z rows): three block
// barriers, B1 after the flow pull loads, B2 after the flow push stage and the Cahn-Hilliard
// pull loads, B3 after the Cahn-Hilliard push stage, which reuses the flow pull region (last
// read before B2).  Blocks (z stage) or waves without face cells take the body without face
// rules (FACES false).
template <class R, bool ODD, bool ZS = false>
__global__ void __launch_bounds__(256) k_mp_step(const MpArgs m) {
    constexpr bool Z = ODD && ZS;
    constexpr int ZQ = Z ? nzdirs<19>() : 1;
    __shared__ R zrf[ZQ][256], zwf[ZQ][256], zrh[ZQ][256];
    __shared__ uint32_t zmf[Z ? 256 : 1], zmh[Z ? 256 : 1];
    const StepArgs& a = m.a;
    const uint32_t tid = threadIdx.x;
    bool slow;
    {
        const uint32_t t = thread_cell(a);
        Cell e;
        locate_at(a, t < a.N ? t : a.N - 1u, e);
        const int face = (e.bits & (uint32_t)(a.f.npmask | a.g.npmask)) != 0u;
        if constexpr (Z)
            slow = __syncthreads_or(face || t >= a.N) != 0;
        else
            slow = __ballot(face) != 0ull;
    }
    auto body = [&](auto faces_tag) {
    constexpr bool FACES = decltype(faces_tag)::value;
    Cell c;
    bool act = true;
    if constexpr (Z) {
        act = locate_clamped(a, c);
        if constexpr (!FACES) act = true;  // no thread past the last cell
    } else {
        if (!locate(a, c)) return;
    }
    const uint32_t S = a.S;
    R P[19];
    mp_stencil<R>(m, c, (const R*)m.phi, P);
    R g[3];
    mp_gradient<R>(P, g);
    const R ph = P[0];
    const R muc = *slot_at((const R*)m.mu, c.id);
    R f[19];
    uint32_t own_f = 0;
    if constexpr (Z) {
        own_f = zpull_issue<19, R, FACES>(f, a.f, c, a, zrf, tid);
        __syncthreads();  // B1
        zpull_finish<19, R>(f, c, a, zrf, tid, own_f);
    } else {
        read_pops<19, ODD, R, FACES>(f, a.f, c, a, kKindFlow);
    }
    R rho, u[3], p, F[3], gr[3];
    mp_macro<R>(m, f, ph, muc, g, rho, u, p, F, gr);
    const R pc = ph < (R)0 ? (R)0 : (ph > (R)1 ? (R)1 : ph);
    const R nu = mp_nu<R>(m, pc);
    const R itf = (R)1 / ((R)3 * nu + (R)0.5);  // 1 / tau_f
    const R kf = (R)1 - (R)0.5 * itf;
    const R uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
#pragma unroll
    for (int i = 0; i < 19; i++) {
        const R w = (R)wgt<19>(i);
        const R cu = cx(i) * u[0] + cy(i) * u[1] + cz(i) * u[2];
        const R cg = cx(i) * gr[0] + cy(i) * gr[1] + cz(i) * gr[2];
        const R cF = cx(i) * F[0] + cy(i) * F[1] + cz(i) * F[2];
        const R si = w * ((R)3 * cu + (R)4.5 * cu * cu - (R)1.5 * uu);
        const R feq = (i == 0 ? (w - (R)1) : w) * (R)3 * p + rho * si;
        const R Fi = kf * w * (R)3 * (cF + cu * cg);
        f[i] = f[i] - itf * (f[i] - feq) + Fi;
    }
    FaceState<R> fs;
    if constexpr (Z)
        zpush_stage<19, R, FACES>(f, a.f, c, a, kKindFlow, zwf, zmf, tid, act, fs);
    else
        write_pops<19, ODD, R, FACES>(f, a.f, c, a, kKindFlow, fs);

    // Cahn-Hilliard lattice
    R* pu = (R*)m.phiu;
    R dpu[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        R* q = slot_at(pu + (size_t)d * S, c.id);
        const R v = ph * u[d];
        dpu[d] = v - *q;
        if (act) *q = v;
    }
    R h[19];
    uint32_t own_h = 0;
    if constexpr (Z) {
        own_h = zpull_issue<19, R, FACES>(h, a.g, c, a, zrh, tid);
        __syncthreads();  // B2
        zpush_finish<19, R, FACES>(a.f, c, a, zwf, zmf, tid);
        zpull_finish<19, R>(h, c, a, zrh, tid, own_h);
    } else {
        read_pops<19, ODD, R, FACES>(h, a.g, c, a, kKindScalar);
    }
    const R ith = (R)(1.0 / m.tau_h), kh = (R)(1.0 - 0.5 / m.tau_h);
    const R em = (R)m.eta * muc;
#pragma unroll
    for (int i = 0; i < 19; i++) {
        const R w = (R)wgt<19>(i);
        const R cu = cx(i) * u[0] + cy(i) * u[1] + cz(i) * u[2];
        const R cd = cx(i) * dpu[0] + cy(i) * dpu[1] + cz(i) * dpu[2];
        const R heq = i == 0 ? ph + (w - (R)1) * em : w * (em + (R)3 * ph * cu);
        h[i] = h[i] - ith * (h[i] - heq) + kh * w * (R)3 * cd;
    }
    if constexpr (Z) {
        zpush_stage<19, R, FACES>(h, a.g, c, a, kKindScalar, zrf, zmh, tid, act,
                                  FaceState<R>());
        __syncthreads();  // B3
        zpush_finish<19, R, FACES>(a.g, c, a, zrf, zmh, tid);
    } else {
        write_pops<19, ODD, R, FACES>(h, a.g, c, a, kKindScalar);
    }
    };
    if (slow)
        body(std::true_type{});
    else
        body(std::false_type{});
}

// p and u (3 x N cell-interleaved, doubles) of the current pre-collision state (phi and mu
// freshly computed from it).
template <class R>
__global__ void __launch_bounds__(256) k_mp_macro(const MpArgs m, double* p_out, double* u_out) {
    Cell c;
    if (!locate(m.a, c)) return;
    R P[19];
    mp_stencil<R>(m, c, (const R*)m.phi, P);
    R g[3];
    mp_gradient<R>(P, g);
    R f[19];
    read_pops_rt<19, R>(f, m.a.f, c, m.a, kKindFlow);
    R rho, u[3], p, F[3], gr[3];
    mp_macro<R>(m, f, P[0], *slot_at((const R*)m.mu, c.id), g, rho, u, p, F, gr);
    p_out[c.id] = (double)p;
    for (int d = 0; d < 3; d++) u_out[3 * (size_t)c.id + d] = (double)u[d];
}

// Start: h at rest equilibrium w_i phi0 (own-slot layout), for the first phi / mu passes.
template <class R>
__global__ void __launch_bounds__(256) k_mp_init_h(const MpArgs m, const double* phi0) {
    Cell c;
    if (!locate(m.a, c)) return;
    R* hb = (R*)m.a.g.pop;
#pragma unroll
    for (int i = 0; i < 19; i++) *pop_at(hb, i, c.id, m.a, m.a.g) = (R)(wgt<19>(i) * phi0[c.id]);
}

// Equilibria of both lattices from phi0, u0 (3N or null), p0 (N or null) and mu of phi0.
template <class R>
__global__ void __launch_bounds__(256) k_mp_init_eq(const MpArgs m, const double* phi0,
                                                    const double* u0, const double* p0) {
    Cell c;
    if (!locate(m.a, c)) return;
    const uint32_t S = m.a.S;
    double u[3] = {0, 0, 0};
    if (u0)
        for (int d = 0; d < 3; d++) u[d] = u0[3 * (size_t)c.id + d];
    const double ph = phi0[c.id], p = p0 ? p0[c.id] : 0.0;
    const double rho = m.rho_g + ph * (m.rho_l - m.rho_g);
    const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double em = m.eta * (double)*slot_at((const R*)m.mu, c.id);
    R* fb = (R*)m.a.f.pop;
    R* hb = (R*)m.a.g.pop;
#pragma unroll
    for (int i = 0; i < 19; i++) {
        const double w = wgt<19>(i);
        const double cu = cx(i) * u[0] + cy(i) * u[1] + cz(i) * u[2];
        const double si = w * (3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
        *pop_at(fb, i, c.id, m.a, m.a.f) = (R)((i == 0 ? w - 1.0 : w) * 3.0 * p + rho * si);
        *pop_at(hb, i, c.id, m.a, m.a.g) =
            (R)(i == 0 ? ph + (w - 1.0) * em : w * (em + 3.0 * ph * cu));
    }
    R* pu = (R*)m.phiu;
    for (int d = 0; d < 3; d++) *slot_at(pu + (size_t)d * S, c.id) = (R)(ph * u[d]);
}

}  // namespace mkg

using namespace mkg;

struct mkg_mp_ctx {
    mkg_mp_params p;
    int fp32 = 0;
    size_t esz = 8;
    uint32_t N = 0, S = 0, F = 0, nblk = 0;
    hipStream_t stream = nullptr;
    // the velocity-pressure lattice f (block lattice 0) and the Cahn-Hilliard lattice h
    // (lattice 1) in one population block
    void* pops = nullptr;
    PopBlock pb;
    void* phi = nullptr;
    void* mu = nullptr;
    void* phiu = nullptr;
    int8_t* flag = nullptr;
    double* scratch = nullptr;  // 19 N doubles
    double* partials = nullptr;
    double* red_out = nullptr;
    int lay = 0;
    int zstage = 0;  // O steps take the z shift through LDS (blocks hold whole z rows)
    int xcd_chunk = 0;  // block order of the launches, fixed at creation
    FaceTab* ftab = nullptr;  // boundary table of the face rules
};

namespace {

int pfail(const std::string& msg) {
    set_error(msg);
    return -1;
}

#define PCHK(expr)                                                                     \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return pfail(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

MpArgs mp_args(const mkg_mp_ctx* h) {
    MpArgs m;
    std::memset(&m, 0, sizeof(m));
    const mkg_mp_params& p = h->p;
    StepArgs& a = m.a;
    a.nx = p.nx;
    a.ny = p.ny;
    a.nz = p.nz;
    a.N = h->N;
    a.S = h->S;
    a.nblk = h->nblk;
    a.F = h->F;
    a.flag = h->flag;
    a.xcd_chunk = h->xcd_chunk;
    a.rho0 = 0.0;  // no moving-wall term at bounce-back faces
    int np = 0;
    for (int f = 0; f < 6; f++) {
        a.f.bc[f] = a.g.bc[f] = p.bc[f];
        if (p.bc[f] != MKG_BC_PERIODIC) np |= 1 << f;
    }
    a.f.npmask = a.g.npmask = np;
    a.f.lay = a.g.lay = h->lay;
    a.f.adv = a.g.adv = 1;
    h->pb.apply(a);
    h->pb.lattice(a.f, h->pops, 0, h->esz);
    h->pb.lattice(a.g, h->pops, 1, h->esz);
    a.f.fid = 0;
    a.g.fid = 1;
    a.ftab = h->ftab;
    a.zstage = h->zstage;
    m.rho_l = p.rho_l;
    m.rho_g = p.rho_g;
    m.nu_l = p.nu_l;
    m.nu_g = p.nu_g;
    m.beta = 12.0 * p.sigma / p.width;
    m.kappa = 1.5 * p.sigma * p.width;
    m.eta = p.mobility / ((1.0 / 3.0) * (p.tau_h - 0.5));
    m.tau_h = p.tau_h;
    for (int d = 0; d < 3; d++) {
        m.gravity[d] = p.gravity[d];
        m.force[d] = p.force[d];
    }
    m.phi = h->phi;
    m.mu = h->mu;
    m.phiu = h->phiu;
    m.wet_mask = (uint32_t)p.wet_faces;
    m.visc = p.visc_model;
    m.wet_k = p.wet_faces ? 4.0 / p.width * std::cos(p.contact_angle * (M_PI / 180.0)) : 0.0;
    return m;
}

template <class R>
void launch_phi_mu(const MpArgs& m, int odd, hipStream_t s) {
    dim3 grid(m.a.nblk), block(256);
    const bool per = !m.a.has_walls && m.a.g.npmask == 0;
    if (odd && m.a.zstage) {
        if (per) k_mp_phi<R, true, true, true><<<grid, block, 0, s>>>(m);
        else k_mp_phi<R, true, true><<<grid, block, 0, s>>>(m);
    } else if (odd) {
        k_mp_phi<R, true><<<grid, block, 0, s>>>(m);
    } else {
        k_mp_phi<R, false><<<grid, block, 0, s>>>(m);
    }
    if (m.a.zstage) k_mp_mu_z<R><<<grid, block, 0, s>>>(m);
    else k_mp_mu<R><<<grid, block, 0, s>>>(m);
}

void mp_destroy_impl(mkg_mp_ctx* h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void* bufs[] = {h->pops, h->phi, h->mu, h->phiu, h->flag, h->scratch, h->partials,
                    h->red_out, h->ftab};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

}  // namespace

extern "C" {

size_t mkg_mp_params_size(void) { return sizeof(mkg_mp_params); }

int mkg_mp_params_default(mkg_mp_params* m) {
    if (!m) return pfail("params is null");
    std::memset(m, 0, sizeof(*m));
    m->nx = m->ny = m->nz = 1;
    m->rho_l = m->rho_g = 1.0;
    m->nu_l = m->nu_g = 0.1;
    m->sigma = 0.01;
    m->width = 4.0;
    m->mobility = 0.1;
    m->tau_h = 1.0;
    m->contact_angle = 90.0;
    return 0;
}

int mkg_mp_create(const mkg_mp_params* p, mkg_mp_handle* out) {
    if (!out) return pfail("out is null");
    *out = nullptr;
    if (!p) return pfail("params is null");
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) return pfail("grid dimensions must be >= 1");
    if (p->precision != MKG_PREC_FP64 && p->precision != MKG_PREC_FP32)
        return pfail("precision must be MKG_PREC_FP64 or MKG_PREC_FP32");
    if (!(p->rho_l > 0) || !(p->rho_g > 0)) return pfail("densities must be > 0");
    if (!(p->nu_l > 0) || !(p->nu_g > 0)) return pfail("viscosities must be > 0");
    if (!(p->width > 0) || !(p->sigma >= 0)) return pfail("width must be > 0 and sigma >= 0");
    if (!(p->mobility > 0) || !(p->tau_h > 0.5)) return pfail("mobility > 0 and tau_h > 1/2 required");
    for (int f = 0; f < 6; f++)
        if (p->bc[f] != MKG_BC_PERIODIC && p->bc[f] != MKG_BC_BOUNCE)
            return pfail("two-phase faces are periodic or bounce (no-slip wall, zero flux)");
    for (int a = 0; a < 3; a++)
        if ((p->bc[2 * a] == 0) != (p->bc[2 * a + 1] == 0))
            return pfail("a periodic axis needs periodic codes on both faces");
    if (p->wet_faces & ~0x3f) return pfail("wet_faces holds face bits 0..5");
    if (p->visc_model < MKG_MP_VISC_NU_LINEAR || p->visc_model > MKG_MP_VISC_MU_HARMONIC)
        return pfail("visc_model must be one of MKG_MP_VISC_*");
    for (int f = 0; f < 6; f++)
        if ((p->wet_faces >> f & 1) && p->bc[f] != MKG_BC_BOUNCE)
            return pfail("a wetting face must be a bounce (wall) face");
    if (p->wet_faces && !(p->contact_angle > 0.0 && p->contact_angle < 180.0))
        return pfail("contact_angle must lie in (0, 180) degrees");
    const double n = (double)p->nx * p->ny * p->nz;
    const double esz = p->precision == MKG_PREC_FP32 ? 4.0 : 8.0;
    if ((n + 65536.0) * esz >= 4294967296.0) return pfail("grid too large for 32-bit offsets");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return pfail("no HIP device available");
    mkg_mp_ctx* h = new mkg_mp_ctx();
    h->p = *p;
    h->fp32 = p->precision == MKG_PREC_FP32;
    h->esz = h->fp32 ? 4 : 8;
    h->N = (uint32_t)n;
    h->S = ((h->N + 127u) & ~127u) + 1024u;
    h->F = 2u * ((uint32_t)p->ny * p->nz + (uint32_t)p->nx * p->nz + (uint32_t)p->nx * p->ny);
    h->nblk = (h->N + 255u) / 256u;
    h->zstage = (p->nz >= 64 && 256 % p->nz == 0) ? 1 : 0;
    if (const char* env = std::getenv("MKG_ZSTAGE")) h->zstage = h->zstage && std::atoi(env) != 0;
    h->xcd_chunk = xcd_chunk_setting(h->esz);
    auto bad = [&](hipError_t e, const char* what) {
        mp_destroy_impl(h);
        return pfail(std::string(what) + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipStreamCreate(&h->stream)) != hipSuccess) return bad(e, "hipStreamCreate");
    const int q[2] = {19, 19};
    h->pb = plan_pop_block(q, 2, h->N, h->S, h->F, h->esz);
    const size_t lb = h->pb.total * h->esz;
    if ((e = hipMalloc(&h->pops, lb)) != hipSuccess) return bad(e, "hipMalloc");
    if ((e = hipMemsetAsync(h->pops, 0, lb, h->stream)) != hipSuccess) return bad(e, "hipMemsetAsync");
    if ((e = hipMalloc(&h->phi, (size_t)h->S * h->esz)) != hipSuccess) return bad(e, "hipMalloc");
    if ((e = hipMalloc(&h->mu, (size_t)h->S * h->esz)) != hipSuccess) return bad(e, "hipMalloc");
    if ((e = hipMalloc(&h->phiu, (size_t)3 * h->S * h->esz)) != hipSuccess) return bad(e, "hipMalloc");
    if ((e = hipMalloc((void**)&h->flag, h->S)) != hipSuccess) return bad(e, "hipMalloc");
    if ((e = hipMemsetAsync(h->flag, 0, h->S, h->stream)) != hipSuccess) return bad(e, "hipMemsetAsync");
    if ((e = hipMalloc((void**)&h->scratch, (size_t)19 * h->N * sizeof(double))) != hipSuccess)
        return bad(e, "hipMalloc");
    if ((e = hipMalloc((void**)&h->partials, (size_t)(h->nblk + 1) * 4 * sizeof(double))) != hipSuccess)
        return bad(e, "hipMalloc");
    if ((e = hipMalloc((void**)&h->red_out, 8 * sizeof(double))) != hipSuccess) return bad(e, "hipMalloc");
    if ((e = facetab_alloc(&h->ftab)) != hipSuccess) return bad(e, "hipMalloc");
    if ((e = facetab_upload(h->ftab, mp_args(h).a, h->stream)) != hipSuccess)
        return bad(e, "boundary table upload");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return bad(e, "hipStreamSynchronize");
    *out = h;
    return 0;
}

int mkg_mp_destroy(mkg_mp_handle h) {
    mp_destroy_impl(h);
    return 0;
}

int mkg_mp_init(mkg_mp_handle h, const double* phi, const double* u, const double* p) {
    if (!h) return pfail("null handle");
    if (!phi) return pfail("phi is null");
    const size_t N = h->N;
    // device copies of the inputs, in the scratch buffer (19 N doubles >= N + 3 N + N)
    double* dphi = h->scratch;
    double* du = u ? h->scratch + N : nullptr;
    double* dp = p ? h->scratch + 4 * N : nullptr;
    PCHK(hipMemcpyAsync(dphi, phi, N * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if (u) PCHK(hipMemcpyAsync(du, u, 3 * N * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if (p) PCHK(hipMemcpyAsync(dp, p, N * sizeof(double), hipMemcpyHostToDevice, h->stream));
    h->lay = 0;
    MpArgs m = mp_args(h);
    dim3 grid(h->nblk), block(256);
    if (h->fp32) {
        k_mp_init_h<float><<<grid, block, 0, h->stream>>>(m, dphi);
        launch_phi_mu<float>(m, 0, h->stream);
        k_mp_init_eq<float><<<grid, block, 0, h->stream>>>(m, dphi, du, dp);
    } else {
        k_mp_init_h<double><<<grid, block, 0, h->stream>>>(m, dphi);
        launch_phi_mu<double>(m, 0, h->stream);
        k_mp_init_eq<double><<<grid, block, 0, h->stream>>>(m, dphi, du, dp);
    }
    PCHK(hipGetLastError());
    PCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int mkg_mp_step(mkg_mp_handle h, int nsteps) {
    if (!h) return pfail("null handle");
    if (nsteps < 0) return pfail("nsteps must be >= 0");
    dim3 grid(h->nblk), block(256);
    for (int s = 0; s < nsteps; s++) {
        MpArgs m = mp_args(h);
        const int odd = h->lay;
        if (h->fp32) {
            launch_phi_mu<float>(m, odd, h->stream);
            if (odd && h->zstage) k_mp_step<float, true, true><<<grid, block, 0, h->stream>>>(m);
            else if (odd) k_mp_step<float, true><<<grid, block, 0, h->stream>>>(m);
            else k_mp_step<float, false><<<grid, block, 0, h->stream>>>(m);
        } else {
            launch_phi_mu<double>(m, odd, h->stream);
            if (odd && h->zstage) k_mp_step<double, true, true><<<grid, block, 0, h->stream>>>(m);
            else if (odd) k_mp_step<double, true><<<grid, block, 0, h->stream>>>(m);
            else k_mp_step<double, false><<<grid, block, 0, h->stream>>>(m);
        }
        h->lay ^= 1;
    }
    PCHK(hipGetLastError());
    return 0;
}

int mkg_mp_sync(mkg_mp_handle h) {
    if (!h) return pfail("null handle");
    PCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int mkg_mp_get(mkg_mp_handle h, int which, double* host, size_t count) {
    if (!h) return pfail("null handle");
    if (!host) return pfail("host buffer is null");
    const size_t N = h->N;
    MpArgs m = mp_args(h);
    dim3 grid(h->nblk), block(256);
    if (which == MKG_MP_POP_FLOW || which == MKG_MP_POP_PHASE) {
        if (count != 19 * N) return pfail("count does not match the field size");
        const LatArgs& L = which == MKG_MP_POP_FLOW ? m.a.f : m.a.g;
        if (launch_canonical(19, h->fp32, m.a, L, which == MKG_MP_POP_FLOW ? 0 : 1, h->scratch,
                             h->stream))
            return pfail("lattice unavailable");
        PCHK(hipMemcpyAsync(host, h->scratch, 19 * N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        PCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    // macroscopic fields: phi and mu of the current populations first
    if (h->fp32) launch_phi_mu<float>(m, h->lay, h->stream);
    else launch_phi_mu<double>(m, h->lay, h->stream);
    if (which == MKG_MP_PHI || which == MKG_MP_MU) {
        if (count != N) return pfail("count does not match the field size");
        launch_to_double(h->fp32, which == MKG_MP_PHI ? h->phi : h->mu, h->scratch, h->N, h->stream);
        PCHK(hipMemcpyAsync(host, h->scratch, N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        PCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    if (which == MKG_MP_PRESSURE || which == MKG_MP_VEL) {
        if (count != (which == MKG_MP_VEL ? 3 * N : N)) return pfail("count does not match the field size");
        if (h->fp32) k_mp_macro<float><<<grid, block, 0, h->stream>>>(m, h->scratch, h->scratch + N);
        else k_mp_macro<double><<<grid, block, 0, h->stream>>>(m, h->scratch, h->scratch + N);
        PCHK(hipMemcpyAsync(host, which == MKG_MP_PRESSURE ? h->scratch : h->scratch + N,
                            count * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        PCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    return pfail("unknown two-phase field id");
}

int mkg_mp_set(mkg_mp_handle h, int which, const double* host, size_t count) {
    if (!h) return pfail("null handle");
    if (!host) return pfail("host buffer is null");
    const size_t N = h->N;
    if (which == MKG_MP_POP_FLOW || which == MKG_MP_POP_PHASE) {
        if (count != 19 * N) return pfail("count does not match the field size");
        // both lattices share one layout: bring the other one to the own-slot layout too
        MpArgs m = mp_args(h);
        if (h->lay) {
            const LatArgs& O = which == MKG_MP_POP_FLOW ? m.a.g : m.a.f;
            const int kind = which == MKG_MP_POP_FLOW ? 1 : 0;
            if (launch_canonical(19, h->fp32, m.a, O, kind, h->scratch, h->stream) ||
                launch_store_pops(19, h->fp32, m.a, O, h->scratch, h->stream))
                return pfail("lattice unavailable");
        }
        PCHK(hipMemcpyAsync(h->scratch, host, 19 * N * sizeof(double), hipMemcpyHostToDevice, h->stream));
        const LatArgs& L = which == MKG_MP_POP_FLOW ? m.a.f : m.a.g;
        if (launch_store_pops(19, h->fp32, m.a, L, h->scratch, h->stream)) return pfail("lattice unavailable");
        PCHK(hipStreamSynchronize(h->stream));
        h->lay = 0;
        return 0;
    }
    if (which == MKG_MP_PHIU) {
        if (count != 3 * N) return pfail("count does not match the field size");
        // cell-interleaved host data to three planes
        for (int d = 0; d < 3; d++) {
            PCHK(hipMemcpy2DAsync(h->scratch, sizeof(double), host + d, 3 * sizeof(double),
                                  sizeof(double), N, hipMemcpyHostToDevice, h->stream));
            launch_from_double(h->fp32, h->scratch, (char*)h->phiu + (size_t)d * h->S * h->esz,
                               h->N, h->stream);
        }
        PCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    return pfail("set takes MKG_MP_POP_FLOW, MKG_MP_POP_PHASE or MKG_MP_PHIU");
}

int mkg_mp_sums(mkg_mp_handle h, double* out) {
    if (!h) return pfail("null handle");
    if (!out) return pfail("out is null");
    MpArgs m = mp_args(h);
    if (h->fp32) launch_phi_mu<float>(m, h->lay, h->stream);
    else launch_phi_mu<double>(m, h->lay, h->stream);
    launch_to_double(h->fp32, h->phi, h->scratch, h->N, h->stream);
    int nout = 0;
    launch_reduce(4, m.a, nullptr, nullptr, nullptr, nullptr, h->scratch, h->partials, h->red_out,
                  &nout, h->stream);
    PCHK(hipMemcpyAsync(out, h->red_out, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    PCHK(hipStreamSynchronize(h->stream));
    return 0;
}

}  // extern "C"
