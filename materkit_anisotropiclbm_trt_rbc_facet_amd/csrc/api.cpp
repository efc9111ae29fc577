// C-ABI layer of the MI355X time step (see include/mkg.h for the contract and the reference
// interfaces each entry point replaces).  Owns device memory and the HIP stream of a handle,
// tracks the AA layout of every lattice, and dispatches the fused kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mkg.h"
#include "kernels.h"
#include "lattice.h"

using namespace mkg;

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
    g_err = msg;
    return -1;
}

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(std::string(#expr) + ": " + hipGetErrorString(e_));            \
    } while (0)

enum { LF = 0, LG = 1, LH = 2, LC = 3 };
// Elements added to the stride S of the per-cell arrays, which is also the slot stride of the
// slot-major population layout.  Measured on the fused 256^3 FP32 step with one array per
// slot (k_step, two boxes): pad 0 gives 0.689 of HBM peak, 1024 gives 0.701-0.707 (256:
// 0.700, 512: 0.692, 2048: 0.696, 4096: 0.693).
constexpr uint32_t kSlotPad = 1024;
// Cells per page of the paged population layout (2^kPageShift).  A 45-slot in-place update at
// 256^3 (FP32) streams at 6.6 TB/s with pages of 2048-16384 cells holding every slot of every
// lattice, 5.7-5.8 TB/s with one array per slot at any pad, 6.3 with pages of 1024 cells and
// 6.3 with 2048-cell pages per lattice.  The fused step at 256^3 measured fastest with pages of
// 8192 cells (k_step 1.011 ms; 1024: 1.033, 2048: 1.027, 4096: 1.020, 16384: 1.025; one array
// per slot: 1.074).
constexpr uint32_t kPageShift = 13;
// Runs of consecutive blocks per XCD (remap_block); 0 keeps the hardware order.  Measured at
// 256^3: FP32 steps are fastest with runs of 16, FP64 steps in hardware order.
constexpr int kXcdChunkF32 = 16;
constexpr int kXcdChunkF64 = 0;

}  // namespace

namespace mkg {
void set_error(const std::string& msg) { g_err = msg; }

FaceTab facetab_of(const StepArgs& a) {
    FaceTab t;
    std::memset(&t, 0, sizeof(t));
    const LatArgs* lat[4] = {&a.f, &a.g, &a.h, &a.c};
    for (int k = 0; k < 4; k++)
        for (int f = 0; f < 6; f++) {
            t.bc[k][f] = lat[k]->bc[f];
            t.fval[k][f] = lat[k]->fval[f];
            t.ref[k] = lat[k]->ref;
        }
    std::memcpy(t.u_face, a.u_face, sizeof(t.u_face));
    t.rho0 = a.rho0;
    t.wall_rot = a.wall_rot;
    std::memcpy(t.wall_omega, a.wall_omega, sizeof(t.wall_omega));
    std::memcpy(t.wall_center, a.wall_center, sizeof(t.wall_center));
    return t;
}

hipError_t facetab_alloc(FaceTab** dev) {
    hipError_t e = hipMalloc((void**)dev, sizeof(FaceTab));
    if (e == hipSuccess) e = hipMemset(*dev, 0, sizeof(FaceTab));
    return e;
}

hipError_t facetab_upload(FaceTab* dev, const StepArgs& a, hipStream_t s) {
    const FaceTab t = facetab_of(a);
    hipError_t e = hipMemcpyAsync(dev, &t, sizeof(t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

int xcd_chunk_setting(size_t esz) {
    int v = esz == 4 ? kXcdChunkF32 : kXcdChunkF64;
    if (const char* env = std::getenv("MKG_XCD_CHUNK")) v = std::atoi(env);
    return v < 0 ? 0 : v;
}

PopBlock plan_pop_block(const int* q, int nlat, uint32_t N, uint32_t S, uint32_t F, size_t esz) {
    PopBlock b;
    uint32_t shift = kPageShift;
    if (const char* env = std::getenv("MKG_PAGE_SHIFT")) {
        const int v = std::atoi(env);
        if (v >= 8 && v <= 16) shift = (uint32_t)v;
    }
    const uint32_t B = 1u << shift;
    size_t qt = 0;
    for (int k = 0; k < nlat; k++) qt += (size_t)q[k];
    if (qt == 0) return b;
    const size_t npages = ((size_t)N + B - 1) / B;
    // FP32 kernels address every lattice through one shared page stride (pages hold every
    // lattice); FP64 kernels take a page stride per lattice (pages hold one lattice)
    const bool shared = esz == 4;
    size_t qmax = 0;
    for (int k = 0; k < nlat; k++) qmax = std::max(qmax, (size_t)q[k]);
    const size_t span = npages * (shared ? qt : qmax) * B + qt * F;  // largest offset range
    bool paged = span * esz < ((size_t)1 << 32);
    if (const char* env = std::getenv("MKG_PAGES")) paged = paged && std::atoi(env) != 0;
    size_t slots;
    if (paged) {
        b.paged = 1;
        b.pg_shift = shift;
        b.pg_mask = B - 1u;
        b.pg_stride = shared ? (uint32_t)(qt * B) : 0u;
        b.sl_stride = B;
        slots = npages * qt * B;
    } else {
        b.sl_stride = S;
        slots = qt * S;
    }
    b.total = slots + qt * F;
    size_t off = 0, foff = slots;
    for (int k = 0; k < nlat; k++) {
        if (paged && !shared) {
            // lattice k: npages pages of its q[k] slots, from element off * B on
            b.slot0[k] = off * npages * B;
            b.lat_stride[k] = (uint32_t)((size_t)q[k] * B);
        } else {
            b.slot0[k] = off * b.sl_stride;
            b.lat_stride[k] = b.pg_stride;
        }
        b.face0[k] = foff;
        off += (size_t)q[k];
        foff += (size_t)q[k] * F;
    }
    return b;
}
}  // namespace mkg

struct mkg_ctx {
    mkg_params p;
    int fp32 = 0;
    size_t esz = 8;
    uint32_t N = 0, S = 0, nblk = 0;
    uint32_t F = 0;  // cells on the six faces (stride of the face buffer)
    int zstage = 0;  // O steps take the z shift through LDS (blocks hold whole z rows)
    int xcd_chunk = 0;  // block order of the launches, fixed at creation
    int dev = 0;
    hipStream_t stream = nullptr;
    // Every lattice's populations and face buffers in one allocation (popbuf, pop_total
    // elements); pop[L] is slot 0 of lattice L and face[L] its face buffer.  Paged or
    // slot-major addressing (pg_*, sl_stride; StepArgs).
    void* popbuf = nullptr;
    size_t pop_total = 0;
    int paged = 0;
    uint32_t pg_shift = 31, pg_mask = 0x7FFFFFFFu, pg_stride = 0, sl_stride = 0;
    void* pop[4] = {nullptr, nullptr, nullptr, nullptr};
    void* face[4] = {nullptr, nullptr, nullptr, nullptr};
    uint32_t lstride[4] = {0, 0, 0, 0};  // page stride of each lattice (FP64 kernels)
    int q[4] = {0, 0, 0, 0};
    int lay[4] = {0, 0, 0, 0};
    int64_t steps[4] = {0, 0, 0, 0};
    void* phi[2] = {nullptr, nullptr};
    int64_t pf_count = 0;
    int8_t* flag = nullptr;
    int16_t* grain = nullptr;
    void* kappa = nullptr;
    double* rot = nullptr;
    FaceTab* ftab = nullptr;    // boundary table of the face rules (uploaded with h->p)
    double* scratch = nullptr;  // 27*N doubles, allocated on first use
    void* nee = nullptr;        // Q x F storage-type values of the velocity faces
    double* partials = nullptr;
    double* red_out = nullptr;
    int has_walls = 0;
    int timing = 0;
    std::vector<hipEvent_t> events;
    double t_last[4] = {0, 0, 0, 0};
};

namespace {

int alloc_zero(void** p, size_t bytes, hipStream_t s) {
    if (bytes == 0) {
        *p = nullptr;
        return 0;
    }
    HIPCHK(hipMalloc(p, bytes));
    HIPCHK(hipMemsetAsync(*p, 0, bytes, s));
    return 0;
}

int ensure_scratch(mkg_ctx* h) {
    if (h->scratch) return 0;
    HIPCHK(hipMalloc((void**)&h->scratch, (size_t)27 * h->N * sizeof(double)));
    return 0;
}

bool combo_supported(int fq, int pq, int th, int so) {
    static const int table[][4] = {{19, 0, 0, 0},  {27, 0, 0, 0}, {19, 0, 1, 0}, {27, 0, 1, 0},
                                   {0, 19, 1, 0},  {0, 7, 1, 0},  {19, 19, 1, 0}, {19, 7, 1, 0},
                                   {27, 19, 1, 0}, {19, 19, 0, 1}, {19, 7, 0, 1}, {0, 19, 0, 1},
                                   {0, 0, 1, 0},   {0, 0, 0, 1},  {0, 19, 0, 0},  {0, 7, 0, 0},
                                   {19, 19, 0, 0}};
    for (auto& t : table)
        if (t[0] == fq && t[1] == pq && t[2] == th && t[3] == so) return true;
    return false;
}

void fill_lat(const mkg_ctx* h, int L, const int32_t* bc, double ref, const double* fval,
              LatArgs& out) {
    std::memset(&out, 0, sizeof(out));
    out.pop = h->pop[L];
    out.face = h->face[L];
    out.pg_stride = h->lstride[L];
    int np = 0;
    for (int f = 0; f < 6; f++) {
        out.bc[f] = bc[f];
        if (bc[f] != MKG_BC_PERIODIC) np |= 1 << f;
        out.fval[f] = fval ? fval[f] : 0.0;
    }
    out.npmask = np;
    out.lay = h->lay[L];
    out.adv = 0;
    out.fid = L;
    out.ref = ref;
}

StepArgs make_args(const mkg_ctx* h) {
    StepArgs a;
    std::memset(&a, 0, sizeof(a));
    const mkg_params& p = h->p;
    a.nx = p.nx;
    a.ny = p.ny;
    a.nz = p.nz;
    a.N = h->N;
    a.S = h->S;
    a.pg_shift = h->pg_shift;
    a.pg_mask = h->pg_mask;
    a.pg_stride = h->pg_stride;
    a.sl_stride = h->sl_stride;
    a.nblk = h->nblk;
    a.F = h->F;
    a.zstage = h->zstage;
    a.xcd_chunk = h->xcd_chunk;
    a.incomp = p.flow_incomp ? 1 : 0;
    a.has_walls = h->has_walls;
    a.has_kappa = (p.has_kappa && h->kappa) ? 1 : 0;
    a.ngrains = p.ngrains;
    a.pf_coupling = p.pf_coupling;
    fill_lat(h, LF, p.bc_f, p.ref_f, nullptr, a.f);
    fill_lat(h, LG, p.bc_g, p.ref_g, nullptr, a.g);
    fill_lat(h, LH, p.bc_h, p.ref_h, p.T_face, a.h);
    fill_lat(h, LC, p.bc_c, p.ref_c, p.C_face, a.c);
    a.phi_cur = h->phi[h->pf_count & 1];
    a.phi_prev = h->phi[(h->pf_count + 1) & 1];
    a.flag = h->flag;
    a.grain = h->grain;
    a.kappa = h->kappa;
    a.rot = h->rot;
    a.ftab = h->ftab;
    a.s_nu = p.s_nu;
    a.s_bulk = p.s_bulk;
    a.s_odd = p.s_odd;
    a.s_even = p.s_even;
    a.rho0 = p.rho0;
    for (int k = 0; k < 3; k++) {
        a.force[k] = p.force[k];
        a.buoy[k] = p.buoy[k];
    }
    a.T_ref = p.T_ref;
    std::memcpy(a.u_face, p.u_face, sizeof(a.u_face));
    a.alpha_l = p.alpha_l;
    a.alpha_s = p.alpha_s;
    a.latent_T = p.latent_T;
    a.D_l = p.D_l;
    a.D_s = p.D_s;
    a.k_part = p.k_part;
    a.latent_C = p.latent_C;
    a.q_min = p.q_min;
    a.at_w = p.at_coef * std::sqrt(p.pf_wdx2 > 0 ? p.pf_wdx2 : 0.0);
    a.pf_dt = p.pf_dt;
    a.pf_wdx2 = p.pf_wdx2;
    a.eps4 = p.eps4;
    a.lambda = p.lambda;
    a.phi_s = p.phi_s;
    a.aniso_kind = p.aniso_kind;
    a.aniso[0] = p.aniso[0];
    a.aniso[1] = p.aniso[1];
    const double cs2g = (p.phase_q == 7) ? 0.25 : 1.0 / 3.0;
    a.omega_g = 1.0 / (0.5 + p.pf_dt * p.pf_wdx2 / cs2g);
    a.pf_model = p.pf_model;
    a.ac_zc = p.pf_wdx2 > 0 ? (1.0 - 0.5 * a.omega_g) / (std::sqrt(2.0) * std::sqrt(p.pf_wdx2)) : 0.0;
    for (int k = 0; k < 3; k++) {
        a.wall_omega[k] = p.wall_omega[k];
        a.wall_center[k] = p.wall_center[k];
    }
    a.wall_rot = (p.wall_omega[0] != 0.0 || p.wall_omega[1] != 0.0 || p.wall_omega[2] != 0.0) ? 1 : 0;
    a.nee_mask = 0;
    for (int f = 0; f < 6; f++)
        if (p.bc_f[f] == MKG_BC_VELOCITY) a.nee_mask |= 1 << f;
    a.nee = h->nee;
    return a;
}

bool has_velocity_faces(const mkg_params& p) {
    for (int f = 0; f < 6; f++)
        if (p.bc_f[f] == MKG_BC_VELOCITY) return true;
    return false;
}

// Buffer of the velocity faces' replaced populations (Q x F), when the flow has such faces.
int ensure_nee(mkg_ctx* h, const mkg_params& p) {
    if (h->nee || !p.flow_q || !has_velocity_faces(p)) return 0;
    HIPCHK(hipMalloc(&h->nee, (size_t)p.flow_q * h->F * h->esz));
    return 0;
}

int alloc_pops(mkg_ctx* h) {
    const PopBlock b = plan_pop_block(h->q, 4, h->N, h->S, h->F, h->esz);
    if (b.total == 0) return 0;
    h->paged = b.paged;
    h->pg_shift = b.pg_shift;
    h->pg_mask = b.pg_mask;
    h->pg_stride = b.pg_stride;
    h->sl_stride = b.sl_stride;
    h->pop_total = b.total;
    if (alloc_zero(&h->popbuf, h->pop_total * h->esz, h->stream)) return -1;
    for (int L = 0; L < 4; L++) {
        if (!h->q[L]) continue;
        h->pop[L] = (char*)h->popbuf + b.slot0[L] * h->esz;
        h->face[L] = (char*)h->popbuf + b.face0[L] * h->esz;
        h->lstride[L] = b.lat_stride[L];
    }
    return 0;
}

int kind_of(int L) { return L == LF ? 0 : 1; }

const int32_t* bc_of(const mkg_params& p, int L) {
    return L == LF ? p.bc_f : (L == LG ? p.bc_g : (L == LH ? p.bc_h : p.bc_c));
}

// Rewrite one lattice into the own-slot layout (lay 0).
int normalize(mkg_ctx* h, int L) {
    if (!h->pop[L] || h->lay[L] == 0) return 0;
    if (ensure_scratch(h)) return -1;
    StepArgs a = make_args(h);
    LatArgs* la = L == LF ? &a.f : (L == LG ? &a.g : (L == LH ? &a.h : &a.c));
    if (launch_canonical(h->q[L], h->fp32, a, *la, kind_of(L), h->scratch, h->stream))
        return fail("normalize: bad lattice");
    HIPCHK(hipGetLastError());
    if (launch_store_pops(h->q[L], h->fp32, a, *la, h->scratch, h->stream))
        return fail("normalize: bad lattice");
    HIPCHK(hipGetLastError());
    h->lay[L] = 0;
    return 0;
}

int validate(const mkg_params* p) {
    if (!p) return fail("params is null");
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) return fail("grid dimensions must be >= 1");
    double n = (double)p->nx * p->ny * p->nz;
    if (n > 2.0e9) return fail("grid too large for 32-bit cell indices");
    // kernels address a slot array with 32-bit byte offsets
    const double esz = p->precision == MKG_PREC_FP32 ? 4.0 : 8.0;
    if ((n + 65536.0) * esz >= 4294967296.0)
        return fail("grid too large: one population slot (cells x element size) must stay below 4 GiB");
    if (p->flow_q != 0 && p->flow_q != 19 && p->flow_q != 27)
        return fail("flow_q must be 0, 19 or 27");
    if (p->phase_q != 0 && p->phase_q != 7 && p->phase_q != 19)
        return fail("phase_q must be 0, 7 or 19");
    if (p->precision != MKG_PREC_FP64 && p->precision != MKG_PREC_FP32)
        return fail("precision must be MKG_PREC_FP64 or MKG_PREC_FP32");
    if (p->ngrains < 0 || p->ngrains > MKG_MAX_GRAINS) return fail("ngrains out of range");
    if (p->pf_model != MKG_PF_KARMA_RAPPEL && p->pf_model != MKG_PF_ALLEN_CAHN)
        return fail("pf_model must be MKG_PF_KARMA_RAPPEL or MKG_PF_ALLEN_CAHN");
    if (p->aniso_kind < MKG_ANISO_CUBIC || p->aniso_kind > MKG_ANISO_FACET)
        return fail("aniso_kind must be one of MKG_ANISO_*");
    if (p->aniso_kind == MKG_ANISO_FACET && !(p->aniso[1] > 0))
        return fail("the facet anisotropy needs a rounding nu > 0");
    if (!combo_supported(p->flow_q, p->phase_q, p->thermal ? 1 : 0, p->solute ? 1 : 0))
        return fail("unsupported combination of lattices (flow_q, phase_q, thermal, solute)");
    const int32_t* bcs[4] = {p->bc_f, p->bc_g, p->bc_h, p->bc_c};
    for (int L = 0; L < 4; L++)
        for (int f = 0; f < 6; f++)
            if (bcs[L][f] < 0 || bcs[L][f] > 5) return fail("face boundary code out of range");
    for (int L = 1; L < 4; L++)
        for (int f = 0; f < 6; f++)
            if (bcs[L][f] == MKG_BC_PRESSURE || bcs[L][f] == MKG_BC_VELOCITY)
                return fail("pressure and velocity faces are for the flow lattice");
    {
        const int dims[3] = {p->nx, p->ny, p->nz};
        for (int f = 0; f < 6; f++) {
            if (p->bc_f[f] != MKG_BC_VELOCITY) continue;
            if (dims[f / 2] < 3)
                return fail("a velocity face needs at least three cells along its normal");
            for (int g = 0; g < 6; g++)
                if (g / 2 != f / 2 && p->bc_f[g] == MKG_BC_VELOCITY)
                    return fail("velocity faces must not share an edge");
        }
    }
    for (int a = 0; a < 3; a++) {
        // a face pair must be periodic on both sides or on neither
        for (int L = 0; L < 4; L++)
            if ((bcs[L][2 * a] == 0) != (bcs[L][2 * a + 1] == 0))
                return fail("periodic faces must come in pairs");
    }
    return 0;
}

int copy_out(const double* dsrc, void* host, int dtype, size_t count, hipStream_t s) {
    if (dtype == MKG_DT_F64) {
        HIPCHK(hipMemcpyAsync(host, dsrc, count * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return 0;
    }
    if (dtype == MKG_DT_F32) {
        std::vector<double> tmp(count);
        HIPCHK(hipMemcpyAsync(tmp.data(), dsrc, count * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        float* o = (float*)host;
        for (size_t i = 0; i < count; i++) o[i] = (float)tmp[i];
        return 0;
    }
    return fail("dtype must be MKG_DT_F64 or MKG_DT_F32 for this field");
}

int copy_in(double* ddst, const void* host, int dtype, size_t count, hipStream_t s) {
    if (dtype == MKG_DT_F64) {
        HIPCHK(hipMemcpyAsync(ddst, host, count * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        return 0;
    }
    if (dtype == MKG_DT_F32) {
        std::vector<double> tmp(count);
        const float* i = (const float*)host;
        for (size_t k = 0; k < count; k++) tmp[k] = (double)i[k];
        HIPCHK(hipMemcpyAsync(ddst, tmp.data(), count * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        return 0;
    }
    return fail("dtype must be MKG_DT_F64 or MKG_DT_F32 for this field");
}

// One time step advancing the lattices flagged in adv.  ev (optional) receives three events:
// before the phase moment pass, between the passes, after the fused pass.  mid (optional)
// runs between the two passes (used by the adjoint to save phi^n and the thermal state).
template <class Mid>
int step_once_impl(mkg_ctx* h, const int* adv, hipEvent_t* ev, Mid mid, void* tape = nullptr) {
    Combo k{h->q[LF], h->q[LG], h->q[LH] ? 1 : 0, h->q[LC] ? 1 : 0};
    int lay0 = -1;
    bool mixed = false;
    for (int L = 0; L < 4; L++)
        if (adv[L]) {
            if (lay0 < 0) lay0 = h->lay[L];
            else if (h->lay[L] != lay0) mixed = true;
        }
    if (mixed) {
        for (int L = 0; L < 4; L++)
            if (adv[L] && normalize(h, L)) return -1;
        lay0 = 0;
    }
    if (ev) HIPCHK(hipEventRecord(ev[0], h->stream));
    if (adv[LG]) {
        h->pf_count++;
        StepArgs a = make_args(h);
        if (launch_phase_moments(h->q[LG], h->fp32, lay0, a, h->phi[h->pf_count & 1], h->stream))
            return fail("phase moment kernel unavailable");
    }
    if (mid(h)) return -1;
    if (ev) HIPCHK(hipEventRecord(ev[1], h->stream));
    StepArgs a = make_args(h);
    a.f.adv = adv[LF];
    a.g.adv = adv[LG];
    a.h.adv = adv[LH];
    a.c.adv = adv[LC];
    a.tape = tape;
    if (adv[LF] && a.nee_mask && launch_nee(k, h->fp32, a, h->stream))
        return fail("velocity-face kernel unavailable");
    if (launch_step(k, h->fp32, lay0, a, h->stream))
        return fail("no kernel instantiation for this lattice combination");
    if (ev) HIPCHK(hipEventRecord(ev[2], h->stream));
    for (int L = 0; L < 4; L++)
        if (adv[L]) {
            h->lay[L] ^= 1;
            h->steps[L]++;
        }
    return 0;
}

// Fields a reduction reads (flow, thermal, solute, phase) and the lattices it needs.
int reduce_needs(const mkg_ctx* h, int which, int* need) {
    if (which < 0 || which > MKG_R_TIP) return fail("unknown reduction");
    need[0] = which == MKG_R_MASS || which == MKG_R_NUSSELT || which == MKG_R_MOMENTUM ||
              which == MKG_R_UMAX;
    need[1] = which == MKG_R_ENERGY || which == MKG_R_NUSSELT;
    need[2] = which == MKG_R_SOLUTE;
    need[3] = which == MKG_R_SOLID || which == MKG_R_PHI || which == MKG_R_TIP;
    if (need[0] && !h->q[LF]) return fail("reduction needs the flow lattice");
    if (need[1] && !h->q[LH]) return fail("reduction needs the thermal lattice");
    if (need[2] && !h->q[LC]) return fail("reduction needs the solute lattice");
    if (need[3] && !h->q[LG]) return fail("reduction needs the phase lattice");
    return 0;
}

// Host-side scaling of the reduced values: the solid fraction is a mean, the Nusselt number
// 1 + <u_z T> H / (alpha dT) (z-vertical Rayleigh-Benard, walls at the z faces).
void finish_reduce(const mkg_ctx* h, int which, double* v, double* out) {
    const double N = (double)h->N;
    if (which == MKG_R_SOLID) v[0] /= N;
    if (which == MKG_R_NUSSELT) {
        const double dT = h->p.T_face[4] - h->p.T_face[5];
        const double H = h->p.nz;
        v[0] = 1.0 + (v[0] / N) * H / (h->p.alpha_l * dT);
    }
    const int nout = (which == MKG_R_MOMENTUM || which == MKG_R_TIP) ? 3 : 1;
    for (int k = 0; k < nout; k++) out[k] = v[k];
}

int step_once(mkg_ctx* h, const int* adv, hipEvent_t* ev, void*) {
    return step_once_impl(h, adv, ev, [](mkg_ctx*) { return 0; });
}

}  // namespace

extern "C" {

int mkg_abi_version(void) { return MKG_ABI_VERSION; }
size_t mkg_params_size(void) { return sizeof(mkg_params); }
const char* mkg_last_error(void) { return g_err.c_str(); }

int mkg_params_default(mkg_params* p) {
    if (!p) return fail("params is null");
    std::memset(p, 0, sizeof(*p));
    p->nx = p->ny = p->nz = 1;
    p->rho0 = 1.0;
    p->s_nu = p->s_bulk = p->s_odd = p->s_even = 1.0;
    p->phi_s = 0.99;
    p->q_min = 1e-3;
    p->latent_T = 0.5;
    p->latent_C = 0.5;
    p->k_part = 1.0;
    p->ref_f = 1.0;
    for (int k = 0; k < MKG_MAX_GRAINS; k++)
        for (int i = 0; i < 3; i++) p->rot[k][4 * i] = 1.0;
    return 0;
}

int mkg_create(const mkg_params* p, mkg_handle* out) {
    if (!out) return fail("out is null");
    *out = nullptr;
    if (validate(p)) return -1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail("no HIP device available");
    mkg_ctx* h = new mkg_ctx();
    h->p = *p;
    h->fp32 = p->precision == MKG_PREC_FP32;
    h->esz = h->fp32 ? 4 : 8;
    h->N = (uint32_t)((size_t)p->nx * p->ny * p->nz);
    // S: the cell count rounded up to 128 elements plus a pad, the stride of the per-cell
    // arrays and of the slot-major population layout, so that its slots do not all start at
    // the same offset modulo a large power of two (which would send the Q loads of a cell to
    // the same memory channel).  MKG_SLOT_PAD overrides the pad.
    {
        uint32_t pad = kSlotPad;
        if (const char* env = std::getenv("MKG_SLOT_PAD")) pad = (uint32_t)std::strtoul(env, nullptr, 10);
        h->S = ((h->N + 127u) & ~127u) + ((pad + 127u) & ~127u);
    }
    h->nblk = (h->N + 255u) / 256u;
    h->F = 2u * ((uint32_t)p->ny * p->nz + (uint32_t)p->nx * p->nz + (uint32_t)p->nx * p->ny);
    // z stage for O steps when a 256-thread block holds whole z rows of at least one wave
    // (nz = 64, 128 or 256); MKG_ZSTAGE=0 turns it off
    h->zstage = (p->nz >= 64 && 256 % p->nz == 0) ? 1 : 0;
    if (const char* env = std::getenv("MKG_ZSTAGE")) h->zstage = h->zstage && std::atoi(env) != 0;
    h->xcd_chunk = xcd_chunk_setting(h->esz);
    h->q[LF] = p->flow_q;
    h->q[LG] = p->phase_q;
    h->q[LH] = p->thermal ? 7 : 0;
    h->q[LC] = p->solute ? 7 : 0;
    auto cleanup = [&](int rc) {
        mkg_destroy(h);
        return rc;
    };
    if (hipGetDevice(&h->dev) != hipSuccess) return cleanup(fail("hipGetDevice failed"));
    if (hipStreamCreate(&h->stream) != hipSuccess) return cleanup(fail("hipStreamCreate failed"));
    {
        std::string err;
        if (noalloc_prepare(h->dev, &err)) return cleanup(fail(err));
    }
    if (alloc_pops(h)) return cleanup(-1);
    if (p->phase_q) {
        for (int b = 0; b < 2; b++)
            if (alloc_zero(&h->phi[b], (size_t)h->S * h->esz, h->stream)) return cleanup(-1);
    }
    if (alloc_zero((void**)&h->flag, h->S, h->stream)) return cleanup(-1);
    if (alloc_zero((void**)&h->grain, (size_t)h->S * 2, h->stream)) return cleanup(-1);
    if (p->has_kappa) {
        if (alloc_zero(&h->kappa, (size_t)h->S * h->esz, h->stream)) return cleanup(-1);
        // default conductivity multiplier 1
        std::vector<double> ones(h->N, 1.0);
        double* tmp = nullptr;
        if (hipMalloc((void**)&tmp, h->N * sizeof(double)) != hipSuccess)
            return cleanup(fail("hipMalloc failed"));
        hipMemcpyAsync(tmp, ones.data(), h->N * sizeof(double), hipMemcpyHostToDevice, h->stream);
        launch_from_double(h->fp32, tmp, h->kappa, h->N, h->stream);
        hipStreamSynchronize(h->stream);
        hipFree(tmp);
    }
    if (ensure_nee(h, *p)) return cleanup(-1);
    if (alloc_zero((void**)&h->rot, sizeof(p->rot), h->stream)) return cleanup(-1);
    if (hipMemcpyAsync(h->rot, p->rot, sizeof(p->rot), hipMemcpyHostToDevice, h->stream) !=
        hipSuccess)
        return cleanup(fail("rotation upload failed"));
    if (alloc_zero((void**)&h->partials, (size_t)(h->nblk + 1024) * 4 * sizeof(double),
                   h->stream))
        return cleanup(-1);
    if (alloc_zero((void**)&h->red_out, 16 * sizeof(double), h->stream)) return cleanup(-1);
    if (facetab_alloc(&h->ftab) != hipSuccess) return cleanup(fail("hipMalloc failed"));
    if (facetab_upload(h->ftab, make_args(h), h->stream) != hipSuccess)
        return cleanup(fail("boundary table upload failed"));
    if (hipStreamSynchronize(h->stream) != hipSuccess) return cleanup(fail("init sync failed"));
    *out = h;
    return 0;
}

int mkg_destroy(mkg_handle h) {
    if (!h) return 0;
    if (h->stream) hipStreamSynchronize(h->stream);
    hipFree(h->popbuf);
    hipFree(h->phi[0]);
    hipFree(h->phi[1]);
    hipFree(h->flag);
    hipFree(h->grain);
    hipFree(h->kappa);
    hipFree(h->rot);
    hipFree(h->ftab);
    hipFree(h->scratch);
    hipFree(h->nee);
    hipFree(h->partials);
    hipFree(h->red_out);
    for (auto e : h->events) hipEventDestroy(e);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

int mkg_get_params(mkg_handle h, mkg_params* p) {
    if (!h || !p) return fail("null argument");
    *p = h->p;
    return 0;
}

int mkg_set_params(mkg_handle h, const mkg_params* p) {
    if (!h || !p) return fail("null argument");
    if (validate(p)) return -1;
    const mkg_params& o = h->p;
    if (p->nx != o.nx || p->ny != o.ny || p->nz != o.nz || p->flow_q != o.flow_q ||
        p->phase_q != o.phase_q || p->thermal != o.thermal || p->solute != o.solute ||
        p->precision != o.precision || (p->has_kappa && !h->kappa))
        return fail("set_params cannot change grid, lattices, precision or kappa storage");
    if (p->ref_f != o.ref_f || p->ref_g != o.ref_g || p->ref_h != o.ref_h || p->ref_c != o.ref_c)
        return fail("set_params cannot change deviation references of live populations");
    // populations still on their way through a face were settled with the old face rules:
    // bring every lattice to the own-slot layout before the new parameters apply
    for (int L = 0; L < 4; L++)
        if (normalize(h, L)) return -1;
    if (ensure_nee(h, *p)) return -1;
    h->p = *p;
    HIPCHK(hipMemcpyAsync(h->rot, p->rot, sizeof(p->rot), hipMemcpyHostToDevice, h->stream));
    HIPCHK(facetab_upload(h->ftab, make_args(h), h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int mkg_init_equilibrium(mkg_handle h, const double* rho, const double* vel3, const double* T,
                         const double* C, const double* phi) {
    if (!h) return fail("null handle");
    if (ensure_scratch(h)) return -1;
    const size_t N = h->N;
    double* d = h->scratch;
    double *drho = nullptr, *du = nullptr, *dT = nullptr, *dC = nullptr, *dphi = nullptr;
    if (rho) { drho = d; HIPCHK(hipMemcpyAsync(drho, rho, N * 8, hipMemcpyHostToDevice, h->stream)); }
    if (vel3) { du = d + N; HIPCHK(hipMemcpyAsync(du, vel3, 3 * N * 8, hipMemcpyHostToDevice, h->stream)); }
    if (T) { dT = d + 4 * N; HIPCHK(hipMemcpyAsync(dT, T, N * 8, hipMemcpyHostToDevice, h->stream)); }
    if (C) { dC = d + 5 * N; HIPCHK(hipMemcpyAsync(dC, C, N * 8, hipMemcpyHostToDevice, h->stream)); }
    if (phi) { dphi = d + 6 * N; HIPCHK(hipMemcpyAsync(dphi, phi, N * 8, hipMemcpyHostToDevice, h->stream)); }
    if (phi && !h->p.phase_q) return fail("phi given but the phase lattice is off");
    StepArgs a = make_args(h);
    launch_init_eq(h->fp32, h->q[LF], h->q[LG], a, drho, du, dT, dC, dphi, h->phi[0], h->phi[1],
                   h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    h->lay[LF] = h->lay[LH] = h->lay[LC] = 0;
    if (phi) h->lay[LG] = 0;
    return 0;
}

int mkg_init_seeds(mkg_handle h, int n, const double* centers, const double* radii) {
    if (!h) return fail("null handle");
    if (!h->p.phase_q) return fail("seeds need the phase lattice");
    if (n < 1 || !centers || !radii) return fail("need at least one seed");
    if (ensure_scratch(h)) return -1;
    double* dc = h->scratch + (size_t)h->N;
    double* dr = dc + 3 * (size_t)n;
    HIPCHK(hipMemcpyAsync(dc, centers, 3 * n * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(dr, radii, n * 8, hipMemcpyHostToDevice, h->stream));
    const double width = std::sqrt(2.0) * std::sqrt(h->p.pf_wdx2);
    const double ts = std::atanh(h->p.phi_s) * width;
    StepArgs a = make_args(h);
    launch_init_seeds(h->fp32, h->q[LG], a, n, dc, dr, width, ts, h->phi[0], h->phi[1],
                      h->scratch, h->grain, h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    h->lay[LG] = 0;
    return 0;
}

int mkg_step(mkg_handle h, int nsteps, int mask) {
    if (!h) return fail("null handle");
    if (nsteps < 0) return fail("nsteps must be >= 0");
    int adv[4] = {(mask & MKG_FLOW) && h->q[LF], (mask & MKG_PHASE) && h->q[LG],
                  (mask & MKG_THERMAL) && h->q[LH], (mask & MKG_SOLUTE) && h->q[LC]};
    if (!(adv[0] || adv[1] || adv[2] || adv[3])) return 0;
    size_t need_ev = h->timing ? (size_t)3 * nsteps : 0;
    while (h->events.size() < need_ev) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        h->events.push_back(e);
    }
    for (int n = 0; n < nsteps; n++) {
        hipEvent_t* ev = h->timing ? &h->events[3 * n] : nullptr;
        if (step_once(h, adv, ev, nullptr)) return -1;
    }
    HIPCHK(hipGetLastError());
    if (h->timing && nsteps > 0) {
        double tk[2] = {0, 0};
        HIPCHK(hipStreamSynchronize(h->stream));
        for (int n = 0; n < nsteps; n++) {
            float a = 0, b = 0;
            HIPCHK(hipEventElapsedTime(&a, h->events[3 * n], h->events[3 * n + 1]));
            HIPCHK(hipEventElapsedTime(&b, h->events[3 * n + 1], h->events[3 * n + 2]));
            tk[0] += a;
            tk[1] += b;
        }
        h->t_last[0] = tk[0];
        h->t_last[1] = tk[1];
        h->t_last[2] = tk[0] + tk[1];
        h->t_last[3] = nsteps;
    }
    return 0;
}

int mkg_sync(mkg_handle h) {
    if (!h) return fail("null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int64_t mkg_step_count(mkg_handle h, int bit) {
    if (!h) return -1;
    int L = bit == MKG_FLOW ? LF : bit == MKG_PHASE ? LG : bit == MKG_THERMAL ? LH : bit == MKG_SOLUTE ? LC : -1;
    return L < 0 ? -1 : h->steps[L];
}

int mkg_enable_timing(mkg_handle h, int on) {
    if (!h) return fail("null handle");
    h->timing = on ? 1 : 0;
    return 0;
}

int mkg_last_timing(mkg_handle h, double* out) {
    if (!h || !out) return fail("null argument");
    for (int k = 0; k < 4; k++) out[k] = h->t_last[k];
    return 0;
}

int mkg_get_field(mkg_handle h, int field, void* host, int dtype, size_t count) {
    if (!h || !host) return fail("null argument");
    const size_t N = h->N;
    if (field >= MKG_F_POP_FLOW && field <= MKG_F_POP_SOLUTE) {
        int L = field - MKG_F_POP_FLOW;
        if (!h->pop[L]) return fail("lattice not present");
        if (count != (size_t)h->q[L] * N) return fail("count must be Q*N");
        if (ensure_scratch(h)) return -1;
        StepArgs a = make_args(h);
        LatArgs* la = L == LF ? &a.f : (L == LG ? &a.g : (L == LH ? &a.h : &a.c));
        launch_canonical(h->q[L], h->fp32, a, *la, kind_of(L), h->scratch, h->stream);
        HIPCHK(hipGetLastError());
        return copy_out(h->scratch, host, dtype, count, h->stream);
    }
    if (field == MKG_F_RHO || field == MKG_F_VEL || field == MKG_F_TEMP || field == MKG_F_CONC) {
        if (field == MKG_F_RHO && !h->q[LF]) return fail("flow lattice not present");
        if (field == MKG_F_TEMP && !h->q[LH]) return fail("thermal lattice not present");
        if (field == MKG_F_CONC && !h->q[LC]) return fail("solute lattice not present");
        size_t want = field == MKG_F_VEL ? 3 * N : N;
        if (count != want) return fail("count mismatch");
        if (ensure_scratch(h)) return -1;
        double* d = h->scratch;
        StepArgs a = make_args(h);
        launch_macro(h->q[LF], h->fp32, a, field == MKG_F_RHO ? d : nullptr,
                     field == MKG_F_VEL ? d : nullptr, field == MKG_F_TEMP ? d : nullptr,
                     field == MKG_F_CONC ? d : nullptr, h->stream);
        HIPCHK(hipGetLastError());
        return copy_out(d, host, dtype, count, h->stream);
    }
    if (field == MKG_F_PHI || field == MKG_F_PHI_PREV || field == MKG_F_KAPPA) {
        const void* src;
        if (field == MKG_F_KAPPA) {
            if (!h->kappa) return fail("kappa not allocated (has_kappa = 0)");
            src = h->kappa;
        } else {
            if (!h->q[LG]) return fail("phase lattice not present");
            src = field == MKG_F_PHI ? h->phi[h->pf_count & 1] : h->phi[(h->pf_count + 1) & 1];
        }
        if (count != N) return fail("count must be N");
        if (ensure_scratch(h)) return -1;
        launch_to_double(h->fp32, src, h->scratch, h->N, h->stream);
        HIPCHK(hipGetLastError());
        return copy_out(h->scratch, host, dtype, count, h->stream);
    }
    if (field == MKG_F_FLAG) {
        if (dtype != MKG_DT_I8 || count != N) return fail("flags are N int8");
        HIPCHK(hipMemcpyAsync(host, h->flag, N, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    if (field == MKG_F_GRAIN) {
        if (dtype != MKG_DT_I16 || count != N) return fail("grain ids are N int16");
        HIPCHK(hipMemcpyAsync(host, h->grain, N * 2, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    return fail("unknown field id");
}

int mkg_set_field(mkg_handle h, int field, const void* host, int dtype, size_t count) {
    if (!h || !host) return fail("null argument");
    const size_t N = h->N;
    if (field >= MKG_F_POP_FLOW && field <= MKG_F_POP_SOLUTE) {
        int L = field - MKG_F_POP_FLOW;
        if (!h->pop[L]) return fail("lattice not present");
        if (count != (size_t)h->q[L] * N) return fail("count must be Q*N");
        if (ensure_scratch(h)) return -1;
        if (copy_in(h->scratch, host, dtype, count, h->stream)) return -1;
        StepArgs a = make_args(h);
        LatArgs* la = L == LF ? &a.f : (L == LG ? &a.g : (L == LH ? &a.h : &a.c));
        launch_store_pops(h->q[L], h->fp32, a, *la, h->scratch, h->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(h->stream));
        h->lay[L] = 0;
        return 0;
    }
    if (field == MKG_F_PHI || field == MKG_F_PHI_PREV || field == MKG_F_KAPPA) {
        void* dst;
        if (field == MKG_F_KAPPA) {
            if (!h->kappa) return fail("kappa not allocated (has_kappa = 0)");
            dst = h->kappa;
        } else {
            if (!h->q[LG]) return fail("phase lattice not present");
            dst = field == MKG_F_PHI ? h->phi[h->pf_count & 1] : h->phi[(h->pf_count + 1) & 1];
        }
        if (count != N) return fail("count must be N");
        if (ensure_scratch(h)) return -1;
        if (copy_in(h->scratch, host, dtype, count, h->stream)) return -1;
        launch_from_double(h->fp32, h->scratch, dst, h->N, h->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    if (field == MKG_F_FLAG) {
        if (dtype != MKG_DT_I8 || count != N) return fail("flags are N int8");
        const int8_t* fl = (const int8_t*)host;
        bool walls = false;
        for (size_t i = 0; i < N; i++) {
            if (fl[i] < -2 || fl[i] > 3) return fail("flag value out of range");
            if (fl[i] == MKG_IS_WALLS) walls = true;
        }
        HIPCHK(hipMemcpyAsync(h->flag, host, N, hipMemcpyHostToDevice, h->stream));
        h->has_walls = walls ? 1 : 0;
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    if (field == MKG_F_GRAIN) {
        if (dtype != MKG_DT_I16 || count != N) return fail("grain ids are N int16");
        const int16_t* g = (const int16_t*)host;
        for (size_t i = 0; i < N; i++)
            if (g[i] < 0 || g[i] > h->p.ngrains) return fail("grain id out of range");
        HIPCHK(hipMemcpyAsync(h->grain, host, N * 2, hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    return fail("field cannot be set directly (use mkg_init_equilibrium)");
}

int mkg_reduce(mkg_handle h, int which, double* out) {
    if (!h || !out) return fail("null argument");
    int need[4];
    if (reduce_needs(h, which, need)) return -1;
    if (ensure_scratch(h)) return -1;
    const size_t N = h->N;
    if (which == MKG_R_NONFINITE) {
        double total = 0;
        if (h->popbuf) {
            launch_nonfinite(h->fp32, h->popbuf, h->pop_total, h->partials, h->red_out, h->stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(&total, h->red_out, 8, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
        }
        out[0] = total;
        return 0;
    }
    double* rho = h->scratch;
    double* u = rho + N;
    double* T = rho + 4 * N;
    double* C = rho + 5 * N;
    double* phi = rho + 6 * N;
    const bool need_flow = need[0], need_T = need[1], need_C = need[2], need_phi = need[3];
    StepArgs a = make_args(h);
    if (need_flow || need_T || need_C) {
        launch_macro(h->q[LF], h->fp32, a, need_flow ? rho : nullptr, need_flow ? u : nullptr,
                     need_T ? T : nullptr, need_C ? C : nullptr, h->stream);
        HIPCHK(hipGetLastError());
    }
    if (need_phi) {
        launch_to_double(h->fp32, h->phi[h->pf_count & 1], phi, h->N, h->stream);
        HIPCHK(hipGetLastError());
    }
    int nout = 1;
    launch_reduce(which, a, rho, u, T, C, phi, h->partials, h->red_out, &nout, h->stream);
    HIPCHK(hipGetLastError());
    double v[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(v, h->red_out, 4 * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    finish_reduce(h, which, v, out);
    return 0;
}

int mkg_reduce_noalloc(mkg_handle h, int which, double* out) {
    if (!h || !out) return fail("null argument");
    int need[4];
    if (reduce_needs(h, which, need)) return -1;
    StepArgs a = make_args(h);
    void* pops[4];
    size_t sizes[4];
    int npops = 0;
    if (h->popbuf) {
        pops[npops] = h->popbuf;
        sizes[npops++] = h->pop_total;
    }
    // cell values needed: flow (with the scalars), the scalars alone, or the phase field
    const int macro = need[0] ? 2 : ((need[1] || need[2]) ? 1 : 0);
    double v[4] = {0, 0, 0, 0};
    std::string err;
    if (reduce_noalloc(h->dev, h->q[LF], h->fp32, a, which, macro, pops, sizes, npops,
                       h->stream, v, &err))
        return fail("mkg_reduce_noalloc: " + err);
    if (which == MKG_R_NONFINITE) {
        out[0] = v[0];
        return 0;
    }
    finish_reduce(h, which, v, out);
    return 0;
}

// ---------------------------------------------------------------- checkpoints
static const char kMagic[8] = {'M', 'K', 'G', 'C', 'K', 'P', 'T', '2'};

int mkg_checkpoint_save(mkg_handle h, const char* path) {
    if (!h || !path) return fail("null argument");
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return fail(std::string("cannot open ") + path);
    HIPCHK(hipStreamSynchronize(h->stream));
    auto wr = [&](const void* p, size_t n) { return std::fwrite(p, 1, n, fp) == n; };
    // the population block is stored as laid out in memory: the header names its layout
    uint64_t hdr[6] = {MKG_ABI_VERSION, sizeof(mkg_params), h->N, h->esz, h->pop_total,
                       ((uint64_t)h->paged << 32) | h->pg_shift};
    bool ok = wr(kMagic, 8) && wr(hdr, sizeof(hdr)) && wr(&h->p, sizeof(h->p)) &&
              wr(h->lay, sizeof(h->lay)) && wr(h->steps, sizeof(h->steps)) &&
              wr(&h->pf_count, sizeof(h->pf_count));
    std::vector<char> buf;
    auto dump = [&](const void* dptr, size_t bytes) {
        if (!dptr || bytes == 0) return true;
        buf.resize(bytes);
        if (hipMemcpy(buf.data(), dptr, bytes, hipMemcpyDeviceToHost) != hipSuccess) return false;
        return wr(buf.data(), bytes);
    };
    if (ok) ok = dump(h->popbuf, h->pop_total * h->esz);
    if (ok && h->q[LG]) ok = dump(h->phi[0], h->S * h->esz) && dump(h->phi[1], h->S * h->esz);
    if (ok) ok = dump(h->flag, h->S) && dump(h->grain, (size_t)h->S * 2);
    if (ok && h->kappa) ok = dump(h->kappa, h->S * h->esz);
    std::fclose(fp);
    if (!ok) return fail("checkpoint write failed");
    return 0;
}

int mkg_checkpoint_load(mkg_handle h, const char* path) {
    if (!h || !path) return fail("null argument");
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return fail(std::string("cannot open ") + path);
    auto rd = [&](void* p, size_t n) { return std::fread(p, 1, n, fp) == n; };
    char magic[8];
    uint64_t hdr[6];
    mkg_params p;
    int lay[4];
    int64_t steps[4], pfc;
    bool ok = rd(magic, 8) && std::memcmp(magic, kMagic, 8) == 0 && rd(hdr, sizeof(hdr)) &&
              hdr[0] == MKG_ABI_VERSION && hdr[1] == sizeof(mkg_params) && hdr[2] == h->N &&
              hdr[3] == h->esz && hdr[4] == h->pop_total &&
              hdr[5] == (((uint64_t)h->paged << 32) | h->pg_shift) && rd(&p, sizeof(p)) && rd(lay, sizeof(lay)) &&
              rd(steps, sizeof(steps)) && rd(&pfc, sizeof(pfc));
    if (!ok) {
        std::fclose(fp);
        return fail("checkpoint header does not match this handle");
    }
    if (p.nx != h->p.nx || p.ny != h->p.ny || p.nz != h->p.nz) {
        std::fclose(fp);
        return fail("checkpoint grid does not match this handle");
    }
    if (p.flow_q != h->p.flow_q || p.phase_q != h->p.phase_q || p.thermal != h->p.thermal ||
        p.solute != h->p.solute || p.precision != h->p.precision ||
        (p.has_kappa != 0) != (h->kappa != nullptr)) {
        std::fclose(fp);
        return fail("checkpoint lattices do not match this handle");
    }
    std::vector<char> buf;
    auto load = [&](void* dptr, size_t bytes) {
        if (!dptr || bytes == 0) return true;
        buf.resize(bytes);
        if (!rd(buf.data(), bytes)) return false;
        return hipMemcpy(dptr, buf.data(), bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (ok) ok = load(h->popbuf, h->pop_total * h->esz);
    if (ok && h->q[LG]) ok = load(h->phi[0], h->S * h->esz) && load(h->phi[1], h->S * h->esz);
    // the flags decide which kernel bodies run (wall cells need the body with face rules)
    bool walls = false;
    if (ok) {
        ok = load(h->flag, h->S);
        for (uint32_t x = 0; ok && x < h->N; x++)
            if ((int8_t)buf[x] == MKG_IS_WALLS) walls = true;
    }
    if (ok) ok = load(h->grain, (size_t)h->S * 2);
    if (ok && h->kappa) ok = load(h->kappa, h->S * h->esz);
    std::fclose(fp);
    if (!ok) return fail("checkpoint body truncated");
    h->has_walls = walls ? 1 : 0;
    h->p = p;
    std::memcpy(h->lay, lay, sizeof(lay));
    std::memcpy(h->steps, steps, sizeof(steps));
    h->pf_count = pfc;
    HIPCHK(hipMemcpy(h->rot, p.rot, sizeof(p.rot), hipMemcpyHostToDevice));
    HIPCHK(facetab_upload(h->ftab, make_args(h), h->stream));
    return 0;
}


// ---------------------------------------------------------------- discrete adjoint
int mkg_flow_adjoint_run(mkg_handle h, int nsteps, const double* u_target, double* loss,
                         double* grad) {
    if (!h || !u_target || !loss || !grad) return fail("null argument");
    const int Q = h->q[LF];
    if (!Q || h->q[LG] || h->q[LH] || h->q[LC])
        return fail("the flow adjoint covers the flow lattice alone (no phase, thermal or solute)");
    if (h->p.flow_incomp) return fail("the flow adjoint covers the compressible equilibrium");
    for (int f = 0; f < 6; f++) {
        const int b = h->p.bc_f[f];
        if (b != MKG_BC_PERIODIC && b != MKG_BC_BOUNCE && b != MKG_BC_OUTFLOW)
            return fail("the flow adjoint takes periodic, bounce or outflow flow faces");
    }
    if (nsteps < 1) return fail("nsteps must be >= 1");
    const size_t N = h->N, QN = (size_t)Q * N;
    {
        std::vector<int8_t> fl(N);
        HIPCHK(hipMemcpy(fl.data(), h->flag, N, hipMemcpyDeviceToHost));
        for (size_t x = 0; x < N; x++)
            if (fl[x] != MKG_IS_FLUID && fl[x] != MKG_IS_WALLS)
                return fail("the flow adjoint takes fluid and wall cells only");
    }
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    // tape: the pre-collision populations of every level the step kernel wrote (storage type,
    // Q slots of S); the final level in canonical FP64 for the loss
    const size_t tstep = (size_t)Q * h->S * h->esz;
    const size_t need = (size_t)nsteps * tstep + (QN + 2 * QN + 3 * N) * 8 + (size_t)8 * h->nblk * 8;
    if (need > free_b / 10 * 9) return fail("the flow adjoint tape does not fit in device memory");
    char* tape = nullptr;
    double *fN = nullptr, *lam[2] = {nullptr, nullptr}, *ut = nullptr, *part = nullptr,
           *out = nullptr;
    auto release = [&]() {
        if (tape) (void)hipFree(tape);
        double* bufs[] = {fN, lam[0], lam[1], ut, part, out};
        for (double* b : bufs)
            if (b) (void)hipFree(b);
    };
    hipError_t e = hipSuccess;
    if ((e = hipMalloc(&tape, (size_t)nsteps * tstep)) != hipSuccess ||
        (e = hipMalloc(&fN, QN * 8)) != hipSuccess ||
        (e = hipMalloc(&lam[0], QN * 8)) != hipSuccess || (e = hipMalloc(&lam[1], QN * 8)) != hipSuccess ||
        (e = hipMalloc(&ut, 3 * N * 8)) != hipSuccess ||
        (e = hipMalloc(&part, (size_t)8 * h->nblk * 8)) != hipSuccess ||
        (e = hipMalloc(&out, 8 * 8)) != hipSuccess) {
        release();
        return fail(std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    int rc = 0;
    do {
        if ((e = hipMemcpyAsync(ut, u_target, 3 * N * 8, hipMemcpyHostToDevice, h->stream)) ||
            (e = hipMemsetAsync(part, 0, (size_t)8 * h->nblk * 8, h->stream))) {
            rc = fail(std::string("hip: ") + hipGetErrorString(e));
            break;
        }
        // forward: each step writes its level's pre-collision populations to the tape
        const int adv[4] = {1, 0, 0, 0};
        for (int n = 0; n < nsteps && rc == 0; n++)
            if (step_once_impl(h, adv, nullptr, [](mkg_ctx*) { return 0; },
                               tape + (size_t)n * tstep))
                rc = -1;
        if (rc) break;
        StepArgs a = make_args(h);
        launch_canonical(Q, h->fp32, a, a.f, 0, fN, h->stream);
        launch_fadj_loss(Q, h->fp32, a, fN, ut, lam[0], part, h->stream);
        int cur = 0;
        for (int n = nsteps - 1; n >= 0; n--) {
            launch_fadj_step(Q, h->fp32, a, tape + (size_t)n * tstep, lam[cur], lam[cur ^ 1],
                             part, h->stream);
            cur ^= 1;
        }
        for (int k = 0; k < 8; k++) launch_sum_partials(part + (size_t)k * h->nblk, h->nblk, out + k, h->stream);
        if ((e = hipGetLastError()) != hipSuccess) {
            rc = fail(std::string("flow adjoint kernels: ") + hipGetErrorString(e));
            break;
        }
        double hv[8];
        if ((e = hipMemcpyAsync(hv, out, 8 * 8, hipMemcpyDeviceToHost, h->stream)) ||
            (e = hipStreamSynchronize(h->stream))) {
            rc = fail(std::string("hip: ") + hipGetErrorString(e));
            break;
        }
        *loss = hv[0];
        for (int k = 0; k < 7; k++) grad[k] = hv[1 + k];
    } while (false);
    (void)hipStreamSynchronize(h->stream);
    release();
    return rc;
}

int mkg_adjoint_run(mkg_handle h, int nsteps, int ckpt, const double* target, double* loss,
                    double* dL_deps4, double* dL_dkappa, double* times) {
    if (!h || !target || !loss || !dL_deps4) return fail("null argument");
    if ((h->q[LF] != 0 && h->q[LF] != 19) || h->q[LC] || h->q[LG] != 19 || h->q[LH] != 7)
        return fail("the adjoint covers the D3Q19 phase field + D3Q7 thermal model, with or "
                    "without a D3Q19 flow");
    // coupled model: the flow advects both scalars and feels the buoyancy
    const bool flow = h->q[LF] == 19;
    if (flow) {
        if (h->p.flow_incomp) return fail("the coupled adjoint covers the compressible equilibrium");
        for (int f = 0; f < 6; f++) {
            const int b = h->p.bc_f[f];
            if (b != MKG_BC_PERIODIC && b != MKG_BC_BOUNCE && b != MKG_BC_OUTFLOW)
                return fail("the coupled adjoint takes periodic, bounce or outflow flow faces");
        }
    }
    if (h->has_walls) return fail("the adjoint does not support wall cells");
    if (h->p.aniso_kind != MKG_ANISO_CUBIC || h->p.pf_model != MKG_PF_KARMA_RAPPEL)
        return fail("the adjoint covers the cubic (Karma-Rappel) anisotropy only");
    if (nsteps < 1 || ckpt < 0) return fail("nsteps must be >= 1 and ckpt >= 0");
    if (dL_dkappa && !h->kappa) return fail("dL/dkappa needs has_kappa = 1");
    const size_t N = h->N, S = h->S, e = h->esz;
    // ckpt == 0: tape -- the forward pass keeps the per-step values the reverse sweep needs,
    // so there is no recompute (one segment of nsteps steps)
    const bool tape = ckpt == 0;
    if (tape) ckpt = nsteps;
    const int nseg = (nsteps + ckpt - 1) / ckpt;
    const int adv[4] = {flow ? 1 : 0, 1, 1, 0};
    // ---- device storage
    struct Ck {
        int lay_g, lay_h, lay_f;
        int64_t pf, steps_g, steps_h, steps_f;
    };
    std::vector<Ck> ck(tape ? 1 : nseg + 1);
    // a checkpoint: every population of the handle (the phase, thermal and, in the coupled
    // model, flow lattices), both phase buffers and the cell flags (S bytes, rounded up to
    // whole elements)
    const size_t PP = h->pop_total;
    const size_t ck_elems = PP + 2 * S + (S + e - 1) / e;
    // flow tape of one step (coupled model): the canonical pre-collision populations, FP64
    const size_t fseg_stride = flow ? (size_t)19 * N : 0;
    // tape of one step, written by the forward kernel: phi^n, then the 7 pre-collision thermal
    // populations in own-slot order (stride S)
    const size_t seg_stride = 8 * S;
    const size_t seg_elems = (size_t)ckpt * seg_stride;
    char *ckbuf = nullptr, *segbuf = nullptr, *work = nullptr;
    double *dtarget = nullptr, *dpart = nullptr, *dscal = nullptr, *fseg = nullptr;
    int rc = 0;
    // events: 0-1 forward, 2-4 reverse phase, 5-6 segment tape filled, 7-8 segment tape free;
    // the reverse sweep runs on a stream of its own (default priority) when checkpointing
    hipEvent_t ev[9] = {};
    hipStream_t rs = h->stream;
    auto release = [&]() {
        hipStreamSynchronize(h->stream);
        if (rs != h->stream) {
            hipStreamSynchronize(rs);
            hipStreamDestroy(rs);
            rs = h->stream;
        }
        for (int k = 0; k < 9; k++)
            if (ev[k]) {
                hipEventDestroy(ev[k]);
                ev[k] = nullptr;
            }
        hipFree(ckbuf);
        hipFree(segbuf);
        hipFree(fseg);
        hipFree(work);
        hipFree(dtarget);
        hipFree(dpart);
        hipFree(dscal);
    };
#define ADJCHK(expr)                                                             \
    do {                                                                         \
        hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess) {                                                  \
            release();                                                           \
            return fail(std::string(#expr) + ": " + hipGetErrorString(e_));      \
        }                                                                        \
    } while (0)
    if (!tape) ADJCHK(hipMalloc((void**)&ckbuf, (nseg + 1) * ck_elems * e));
    // two segment tapes when checkpointing: the recompute of the next (earlier) segment fills
    // one while the reverse sweep of the current segment reads the other
    const int nbuf = tape ? 1 : 2;
    ADJCHK(hipMalloc((void**)&segbuf, nbuf * seg_elems * e));
    if (flow) ADJCHK(hipMalloc((void**)&fseg, nbuf * (size_t)ckpt * fseg_stride * sizeof(double)));
    // work: Gin, Gout (19N), nin, nout (N), Hin, Hout (7N), pin, pout (N), P (19N), kap (N);
    // coupled: Fin, Fout (19N), u~ adjoint (3N), thermal cell-wide in / out (N)
    const size_t work_elems =
        ((2 * 19 + 2 + 2 * 7 + 2 + 19 + 1) + (flow ? 2 * 19 + 3 + 2 : 0)) * N;
    ADJCHK(hipMalloc((void**)&work, work_elems * e));
    ADJCHK(hipMemsetAsync(work, 0, work_elems * e, h->stream));
    ADJCHK(hipMalloc((void**)&dtarget, N * sizeof(double)));
    ADJCHK(hipMemcpyAsync(dtarget, target, N * sizeof(double), hipMemcpyHostToDevice, h->stream));
    ADJCHK(hipMalloc((void**)&dpart, (size_t)h->nblk * 2 * sizeof(double)));
    ADJCHK(hipMemsetAsync(dpart, 0, (size_t)h->nblk * 2 * sizeof(double), h->stream));
    ADJCHK(hipMalloc((void**)&dscal, 4 * sizeof(double)));
    char* Gin = work;
    char* Gout = Gin + 19 * N * e;
    char* nin = Gout + 19 * N * e;
    char* nout = nin + N * e;
    char* Hin = nout + N * e;
    char* Hout = Hin + 7 * N * e;
    char* pin = Hout + 7 * N * e;
    char* pout = pin + N * e;
    char* Pb = pout + N * e;
    char* kap = Pb + 19 * N * e;
    char* Fin = kap + N * e;
    char* Fout = Fin + 19 * N * e;
    char* Ub = Fout + 19 * N * e;
    char* Cin = Ub + 3 * N * e;
    char* Cout = Cin + N * e;
    double* eps_part = dpart;
    double* loss_part = dpart + h->nblk;

    auto save_ck = [&](int k) -> int {
        char* base = ckbuf + (size_t)k * ck_elems * e;
        HIPCHK(hipMemcpyAsync(base, h->popbuf, PP * e, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(base + PP * e, h->phi[0], S * e, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(base + (PP + S) * e, h->phi[1], S * e, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(base + (PP + 2 * S) * e, h->flag, S, hipMemcpyDeviceToDevice, h->stream));
        ck[k] = {h->lay[LG], h->lay[LH], h->lay[LF], h->pf_count, h->steps[LG], h->steps[LH],
                 h->steps[LF]};
        return 0;
    };
    auto load_ck = [&](int k) -> int {
        char* base = ckbuf + (size_t)k * ck_elems * e;
        HIPCHK(hipMemcpyAsync(h->popbuf, base, PP * e, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(h->phi[0], base + PP * e, S * e, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(h->phi[1], base + (PP + S) * e, S * e, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(h->flag, base + (PP + 2 * S) * e, S, hipMemcpyDeviceToDevice, h->stream));
        h->lay[LG] = ck[k].lay_g;
        h->lay[LH] = ck[k].lay_h;
        h->pf_count = ck[k].pf;
        h->steps[LG] = ck[k].steps_g;
        h->steps[LH] = ck[k].steps_h;
        if (flow) {
            h->lay[LF] = ck[k].lay_f;
            h->steps[LF] = ck[k].steps_f;
        }
        return 0;
    };
    for (int k = 0; k < 9; k++) ADJCHK(hipEventCreate(&ev[k]));
    if (!tape) ADJCHK(hipStreamCreate(&rs));
    auto seg_base = [&](int sgi) { return segbuf + (size_t)(sgi % nbuf) * seg_elems * e; };
    auto fseg_base = [&](int sgi) { return fseg + (size_t)(sgi % nbuf) * ckpt * fseg_stride; };
    // a forward step that writes its tape slot: phi^n and the thermal populations before it
    // (by the step kernel), and with a flow the canonical flow populations before it
    auto stepper = [&](int n, int n0, char* base, double* fbase) -> int {
        char* slot = base + (size_t)(n - n0) * seg_stride * e;
        if (flow) {
            StepArgs a = make_args(h);
            if (launch_canonical(19, h->fp32, a, a.f, 0, fbase + (size_t)(n - n0) * fseg_stride,
                                 h->stream))
                return fail("flow tape kernel unavailable");
        }
        return step_once_impl(h, adv, nullptr, [](mkg_ctx*) { return 0; }, slot);
    };
    // recompute of segment sgi from its checkpoint on the handle's stream, into its tape
    auto recompute = [&](int sgi) -> int {
        const int n0 = sgi * ckpt, n1 = std::min(nsteps, n0 + ckpt);
        if (load_ck(sgi)) return -1;
        for (int n = n0; n < n1; n++)
            if (stepper(n, n0, seg_base(sgi), fseg_base(sgi))) return -1;
        HIPCHK(hipEventRecord(ev[5 + sgi % 2], h->stream));
        return 0;
    };
    ADJCHK(hipEventRecord(ev[0], h->stream));
    // ---- forward pass: checkpoints, or the tape.  With checkpoints, the last segment writes
    // its tape as it steps (its recompute would repeat exactly these steps, and the reverse
    // sweep starts with it), so it needs no checkpoint of its own; every earlier segment
    // keeps a checkpoint and is recomputed.
    const int last0 = (nseg - 1) * ckpt;
    for (int n = 0; n < nsteps; n++) {
        if (tape) {
            if (stepper(n, 0, segbuf, fseg)) { rc = -1; break; }
            continue;
        }
        if (n >= last0) {
            if (stepper(n, last0, seg_base(nseg - 1), fseg_base(nseg - 1))) { rc = -1; break; }
            continue;
        }
        if (n % ckpt == 0 && save_ck(n / ckpt)) { rc = -1; break; }
        if (step_once(h, adv, nullptr, nullptr)) { rc = -1; break; }
    }
    if (rc == 0 && !tape && save_ck(nseg)) rc = -1;  // final forward state
    if (rc) { release(); return -1; }
    // ---- loss and seed of the reverse sweep (phi_N - target, cell-wide)
    {
        StepArgs a = make_args(h);
        launch_adj_loss(h->fp32, a, dtarget, nin, loss_part, dscal, h->stream);
        ADJCHK(hipGetLastError());
    }
    ADJCHK(hipEventRecord(ev[1], h->stream));
    // ---- reverse sweep, segment by segment from the last.  With checkpoints, segment s - 1 is
    // recomputed on the handle's stream while the reverse of segment s runs on rs; a segment
    // tape is refilled only after the reverse that read it has finished.  The reverse kernels
    // read geometry, faces, rates, kappa and flags, never the handle's populations.
    const StepArgs ar = make_args(h);
    // stencil transpose by column sums when a block holds whole z rows (MKG_ADJ_GROUPED=0
    // selects the per-direction form)
    int grouped = h->zstage;
    if (const char* env = std::getenv("MKG_ADJ_GROUPED")) grouped = grouped && std::atoi(env) != 0;
    // The last segment's tape came with the forward pass: nothing is recomputed before the
    // reverse sweep starts (ev[2]..ev[3] keep the recompute slot of the timing report).  Its
    // "tape filled" event is recorded here, after the loss kernel has written the seed that
    // the reverse sweep reads first.
    ADJCHK(hipEventRecord(ev[2], h->stream));
    if (!tape) ADJCHK(hipEventRecord(ev[5 + (nseg - 1) % 2], h->stream));
    ADJCHK(hipEventRecord(ev[3], h->stream));
    for (int sgi = nseg - 1; sgi >= 0 && rc == 0; sgi--) {
        const int n0 = sgi * ckpt, n1 = std::min(nsteps, n0 + ckpt);
        if (!tape) {
            ADJCHK(hipStreamWaitEvent(rs, ev[5 + sgi % 2], 0));
            if (sgi > 0) {
                if (sgi + 1 <= nseg - 1) ADJCHK(hipStreamWaitEvent(h->stream, ev[7 + (sgi - 1) % 2], 0));
                if (recompute(sgi - 1)) { rc = -1; break; }
            }
        }
        char* base = seg_base(sgi);
        for (int n = n1 - 1; n >= n0; n--) {
            char* slot = base + (size_t)(n - n0) * seg_stride * e;
            AdjArgs j;
            std::memset(&j, 0, sizeof(j));
            j.gin = Gin; j.nin = nin; j.hin = Hin; j.pin = pin;
            j.gout = Gout; j.hout = Hout; j.pout = pout;
            j.P = Pb; j.kap_grad = kap; j.phin = slot;
            j.htape = slot + S * e;
            j.eps_part = eps_part;
            if (flow) {
                // scalars and phase with advection, stencil transpose, then the flow
                j.ftape = fseg_base(sgi) + (size_t)(n - n0) * fseg_stride;
                j.fin = Fin; j.fout = Fout; j.ut = Ub;
                j.hcw_in = Cin; j.hcw_out = Cout; j.nacc = nout;
                launch_adj_collide_flow(h->fp32, grouped, ar, j, rs);
                if (grouped) launch_adj_qsum(h->fp32, ar, Pb, nout, rs);
                else launch_adj_stencil(h->fp32, ar, Pb, nout, rs);
                launch_cadj_flow(h->fp32, ar, j, rs);
                std::swap(Fin, Fout);
                std::swap(Cin, Cout);
            } else {
                launch_adj_collide(h->fp32, grouped, ar, j, rs);
                if (grouped) launch_adj_qsum(h->fp32, ar, Pb, nout, rs);
                else launch_adj_stencil(h->fp32, ar, Pb, nout, rs);
            }
            std::swap(Gin, Gout);
            std::swap(nin, nout);
            std::swap(Hin, Hout);
            std::swap(pin, pout);
        }
        if (!tape) ADJCHK(hipEventRecord(ev[7 + sgi % 2], rs));
        if (hipGetLastError() != hipSuccess) rc = fail("adjoint kernel launch failed");
    }
    if (rc) { release(); return -1; }
    ADJCHK(hipEventRecord(ev[4], rs));
    // ---- outputs
    launch_sum_partials(eps_part, h->nblk, dscal + 1, rs);
    ADJCHK(hipGetLastError());
    double sc[2];
    ADJCHK(hipMemcpyAsync(sc, dscal, 2 * sizeof(double), hipMemcpyDeviceToHost, rs));
    if (dL_dkappa) {
        if (ensure_scratch(h)) { release(); return -1; }
        launch_to_double(h->fp32, kap, h->scratch, h->N, rs);
        ADJCHK(hipMemcpyAsync(dL_dkappa, h->scratch, N * sizeof(double), hipMemcpyDeviceToHost,
                              rs));
    }
    ADJCHK(hipStreamSynchronize(rs));
    float t_re = 0, t_rev = 0;
    ADJCHK(hipEventElapsedTime(&t_re, ev[2], ev[3]));
    ADJCHK(hipEventElapsedTime(&t_rev, ev[3], ev[4]));
    // leave the handle at the final forward state (the tape never moved it)
    if (!tape && load_ck(nseg)) { release(); return -1; }
    ADJCHK(hipStreamSynchronize(h->stream));
    float t_fwd = 0;
    hipEventElapsedTime(&t_fwd, ev[0], ev[1]);
    *loss = sc[0];
    *dL_deps4 = sc[1];
    if (times) {
        times[0] = t_fwd;
        times[1] = t_re;
        times[2] = t_rev;
        times[3] = t_fwd + t_re + t_rev;
    }
    release();
#undef ADJCHK
    return 0;
}

}  // extern "C"
