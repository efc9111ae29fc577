// Auxiliary kernels: layout-independent population read/write, macroscopic fields,
// initial states, seed geometry and the FP64 diagnostic reductions.
#include <cmath>
#include <string>

#include "../../include/mkg.h"
#include "cellred.h"

namespace mkg {

void set_error(const std::string& msg);

int launch_step_f32(const Combo& k, int odd, const StepArgs& a, hipStream_t s);
int launch_step_f64(const Combo& k, int odd, const StepArgs& a, hipStream_t s);
int launch_phase_moments_f32(int pq, int odd, const StepArgs& a, void* out, hipStream_t s);
int launch_phase_moments_f64(int pq, int odd, const StepArgs& a, void* out, hipStream_t s);
int launch_nee_f32(const Combo& k, const StepArgs& a, hipStream_t s);
int launch_nee_f64(const Combo& k, const StepArgs& a, hipStream_t s);

int launch_nee(const Combo& k, int fp32, const StepArgs& a, hipStream_t s) {
    return fp32 ? launch_nee_f32(k, a, s) : launch_nee_f64(k, a, s);
}

int launch_step(const Combo& k, int fp32, int odd, const StepArgs& a, hipStream_t s) {
    return fp32 ? launch_step_f32(k, odd, a, s) : launch_step_f64(k, odd, a, s);
}
int launch_phase_moments(int pq, int fp32, int odd, const StepArgs& a, void* phi_out,
                         hipStream_t s) {
    return fp32 ? launch_phase_moments_f32(pq, odd, a, phi_out, s)
                : launch_phase_moments_f64(pq, odd, a, phi_out, s);
}

static inline dim3 grid_of(const StepArgs& a) { return dim3(a.nblk); }

// ------------------------------------------------------------------ canonical populations
// Wall cells are never touched by a step, so their populations stay in their own slots.
template <class R, int Q>
__global__ void __launch_bounds__(256) k_canonical(const StepArgs a, const LatArgs L, int kind,
                                                   double* __restrict__ out) {
    Cell c;
    if (!locate(a, c)) return;
    R p[Q];
    if (a.flag[c.id] == kIsWalls)
        read_pops<Q, false, R>(p, L, c, a, kind);
    else
        read_pops_rt<Q, R>(p, L, c, a, kind);
#pragma unroll
    for (int i = 0; i < Q; i++) out[(size_t)i * a.N + c.id] = (double)p[i] + wgt<Q>(i) * L.ref;
}

template <class R, int Q>
__global__ void __launch_bounds__(256) k_store(const StepArgs a, const LatArgs L,
                                               const double* __restrict__ in) {
    Cell c;
    if (!locate(a, c)) return;
    R* base = (R*)L.pop;
#pragma unroll
    for (int i = 0; i < Q; i++)
        *pop_at(base, i, c.id, a, L) = (R)(in[(size_t)i * a.N + c.id] - wgt<Q>(i) * L.ref);
}

template <class R>
static int canonical_q(int q, const StepArgs& a, const LatArgs& L, int kind, double* out,
                       hipStream_t s) {
    if (q == 7) k_canonical<R, 7><<<grid_of(a), 256, 0, s>>>(a, L, kind, out);
    else if (q == 19) k_canonical<R, 19><<<grid_of(a), 256, 0, s>>>(a, L, kind, out);
    else if (q == 27) k_canonical<R, 27><<<grid_of(a), 256, 0, s>>>(a, L, kind, out);
    else return -1;
    return 0;
}
template <class R>
static int store_q(int q, const StepArgs& a, const LatArgs& L, const double* in, hipStream_t s) {
    if (q == 7) k_store<R, 7><<<grid_of(a), 256, 0, s>>>(a, L, in);
    else if (q == 19) k_store<R, 19><<<grid_of(a), 256, 0, s>>>(a, L, in);
    else if (q == 27) k_store<R, 27><<<grid_of(a), 256, 0, s>>>(a, L, in);
    else return -1;
    return 0;
}

int launch_canonical(int q, int fp32, const StepArgs& a, const LatArgs& L, int kind,
                     double* out, hipStream_t s) {
    return fp32 ? canonical_q<float>(q, a, L, kind, out, s)
                : canonical_q<double>(q, a, L, kind, out, s);
}
int launch_store_pops(int q, int fp32, const StepArgs& a, const LatArgs& L, const double* in,
                      hipStream_t s) {
    return fp32 ? store_q<float>(q, a, L, in, s) : store_q<double>(q, a, L, in, s);
}

// ------------------------------------------------------------------ macroscopic fields
template <class R, int FQ>
__global__ void __launch_bounds__(256) k_macro(const StepArgs a, double* rho, double* u,
                                               double* T, double* C) {
    Cell c;
    if (!locate(a, c)) return;
    const bool wall = a.flag[c.id] == kIsWalls;
    const CellMacro m = cell_macro<R, FQ>(a, c, wall, C != nullptr);
    if (T && a.h.pop) T[c.id] = m.T;
    if (C && a.c.pop) C[c.id] = m.C;
    if (rho && FQ > 0) rho[c.id] = m.rho;
    if (u)
        for (int k = 0; k < 3; k++) u[3 * (size_t)c.id + k] = m.u[k];
}

template <class R>
static int macro_r(int fq, const StepArgs& a, double* rho, double* u, double* T, double* C,
                   hipStream_t s) {
    if (fq == 0) k_macro<R, 0><<<grid_of(a), 256, 0, s>>>(a, rho, u, T, C);
    else if (fq == 19) k_macro<R, 19><<<grid_of(a), 256, 0, s>>>(a, rho, u, T, C);
    else if (fq == 27) k_macro<R, 27><<<grid_of(a), 256, 0, s>>>(a, rho, u, T, C);
    else return -1;
    return 0;
}
int launch_macro(int fq, int fp32, const StepArgs& a, double* rho, double* u, double* T,
                 double* C, hipStream_t s) {
    return fp32 ? macro_r<float>(fq, a, rho, u, T, C, s) : macro_r<double>(fq, a, rho, u, T, C, s);
}

// ------------------------------------------------------------------ conversions
template <class R>
__global__ void k_to_double(const R* __restrict__ src, double* __restrict__ dst, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) dst[i] = (double)src[i];
}
template <class R>
__global__ void k_from_double(const double* __restrict__ src, R* __restrict__ dst, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) dst[i] = (R)src[i];
}
int launch_to_double(int fp32, const void* src, double* dst, uint32_t n, hipStream_t s) {
    dim3 g((n + 255) / 256);
    if (fp32) k_to_double<float><<<g, 256, 0, s>>>((const float*)src, dst, n);
    else k_to_double<double><<<g, 256, 0, s>>>((const double*)src, dst, n);
    return 0;
}
int launch_from_double(int fp32, const double* src, void* dst, uint32_t n, hipStream_t s) {
    dim3 g((n + 255) / 256);
    if (fp32) k_from_double<float><<<g, 256, 0, s>>>(src, (float*)dst, n);
    else k_from_double<double><<<g, 256, 0, s>>>(src, (double*)dst, n);
    return 0;
}

// ------------------------------------------------------------------ initial states
// Scalar lattices at rest equilibrium w_i * value; phase buffers and flags from phi.
template <class R>
__global__ void __launch_bounds__(256) k_init_eq(const StepArgs a, const double* T,
                                                 const double* C, const double* phi, R* phi0,
                                                 R* phi1) {
    Cell c;
    if (!locate(a, c)) return;
    const uint32_t id = c.id;
    if (a.g.pop && phi) {
        const double ph = phi[id];
        phi0[id] = (R)ph;
        phi1[id] = (R)ph;
        const int8_t fl = a.flag[id];
        const int8_t nf = classify(fl, ph, a.phi_s, a.pf_model);
        if (nf != fl) a.flag[id] = nf;
    }
    if (a.h.pop) {
        const double t = T ? T[id] : a.T_ref;
        R* h = (R*)a.h.pop;
#pragma unroll
        for (int i = 0; i < 7; i++) *pop_at(h, i, id, a, a.h) = (R)(wgt<7>(i) * (t - a.h.ref));
    }
    if (a.c.pop) {
        const double cv = C ? C[id] : 0.0;
        R* p = (R*)a.c.pop;
#pragma unroll
        for (int i = 0; i < 7; i++) *pop_at(p, i, id, a, a.c) = (R)(wgt<7>(i) * (cv - a.c.ref));
    }
}

// Equilibrium populations of a Q lattice from (rho, u): own-slot layout, deviations.
template <class R, int Q>
__global__ void __launch_bounds__(256) k_init_flow(const StepArgs a, const double* rho,
                                                   const double* u) {
    Cell c;
    if (!locate(a, c)) return;
    const uint32_t id = c.id;
    double ux = 0, uy = 0, uz = 0;
    if (u) {
        ux = u[3 * (size_t)id];
        uy = u[3 * (size_t)id + 1];
        uz = u[3 * (size_t)id + 2];
    }
    const double r = rho ? rho[id] : a.rho0;
    const double uu = ux * ux + uy * uy + uz * uz;
    R* f = (R*)a.f.pop;
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const double cu = cx(i) * ux + cy(i) * uy + cz(i) * uz;
        const double feq =
            a.incomp ? wgt<Q>(i) * (r + a.rho0 * (3.0 * cu + 4.5 * cu * cu - 1.5 * uu))
                     : wgt<Q>(i) * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
        *pop_at(f, i, id, a, a.f) = (R)(feq - wgt<Q>(i) * a.f.ref);
    }
}

template <class R, int Q>
__global__ void __launch_bounds__(256) k_init_phase(const StepArgs a, const double* phi) {
    Cell c;
    if (!locate(a, c)) return;
    R* g = (R*)a.g.pop;
    const double ph = phi[c.id];
#pragma unroll
    for (int i = 0; i < Q; i++) *pop_at(g, i, c.id, a, a.g) = (R)(wgt<Q>(i) * (ph - a.g.ref));
}

template <class R>
static int init_eq_r(const StepArgs& a, int fq, int pq, const double* rho, const double* u,
                     const double* T, const double* C, const double* phi, R* phi0, R* phi1,
                     hipStream_t s) {
    k_init_eq<R><<<grid_of(a), 256, 0, s>>>(a, T, C, phi, phi0, phi1);
    if (fq == 19) k_init_flow<R, 19><<<grid_of(a), 256, 0, s>>>(a, rho, u);
    if (fq == 27) k_init_flow<R, 27><<<grid_of(a), 256, 0, s>>>(a, rho, u);
    if (phi && pq == 19) k_init_phase<R, 19><<<grid_of(a), 256, 0, s>>>(a, phi);
    if (phi && pq == 7) k_init_phase<R, 7><<<grid_of(a), 256, 0, s>>>(a, phi);
    return 0;
}

// ------------------------------------------------------------------ seeds
// phi = -tanh(d / width) with d the distance to the nearest seed surface; grain = that seed;
// flags from d against +-ts = atanh(phi_s) * width so that they are exact integers of the
// geometry (mkMultiphaseField::initialize_seeds, MaterKit/inc/mkMultiphaseField.h:52).
template <class R>
__global__ void __launch_bounds__(256) k_seeds(const StepArgs a, int n, const double* centers,
                                               const double* radii, double width, double ts,
                                               R* phi0, R* phi1, double* phid, int16_t* grain) {
    Cell c;
    if (!locate(a, c)) return;
    double best = 1e300;
    int gbest = 0;
    for (int k = 0; k < n; k++) {
        const double dx = c.x - centers[3 * k], dy = c.y - centers[3 * k + 1],
                     dz = c.z - centers[3 * k + 2];
        const double d = sqrt(dx * dx + dy * dy + dz * dz) - radii[k];
        if (d < best) {
            best = d;
            gbest = k + 1;
        }
    }
    const double ph = -tanh(best / width);
    phi0[c.id] = (R)ph;
    phi1[c.id] = (R)ph;
    phid[c.id] = ph;
    if (grain) grain[c.id] = (int16_t)gbest;
    const int8_t fl = a.flag[c.id];
    if (fl != kIsWalls && fl >= 0 && a.pf_model == 0)
        a.flag[c.id] = (int8_t)((best <= -ts) ? kIsSolid : (best < ts ? kIsInter : kIsFluid));
}

int launch_init_eq(int fp32, int fq, int pq, const StepArgs& a, const double* rho,
                   const double* u, const double* T, const double* C, const double* phi,
                   void* phi0, void* phi1, hipStream_t s) {
    if (fp32)
        return init_eq_r<float>(a, fq, pq, rho, u, T, C, phi, (float*)phi0, (float*)phi1, s);
    return init_eq_r<double>(a, fq, pq, rho, u, T, C, phi, (double*)phi0, (double*)phi1, s);
}

int launch_init_seeds(int fp32, int pq, const StepArgs& a, int n, const double* centers,
                      const double* radii, double width, double ts, void* phi0, void* phi1,
                      double* phid, int16_t* grain, hipStream_t s) {
    if (fp32)
        k_seeds<float><<<grid_of(a), 256, 0, s>>>(a, n, centers, radii, width, ts, (float*)phi0,
                                                  (float*)phi1, phid, grain);
    else
        k_seeds<double><<<grid_of(a), 256, 0, s>>>(a, n, centers, radii, width, ts,
                                                   (double*)phi0, (double*)phi1, phid, grain);
    if (fp32) {
        if (pq == 19) k_init_phase<float, 19><<<grid_of(a), 256, 0, s>>>(a, phid);
        if (pq == 7) k_init_phase<float, 7><<<grid_of(a), 256, 0, s>>>(a, phid);
    } else {
        if (pq == 19) k_init_phase<double, 19><<<grid_of(a), 256, 0, s>>>(a, phid);
        if (pq == 7) k_init_phase<double, 7><<<grid_of(a), 256, 0, s>>>(a, phid);
    }
    return 0;
}

// ------------------------------------------------------------------ reductions
__global__ void __launch_bounds__(256) k_reduce(const StepArgs a, int which, const double* rho,
                                                const double* u, const double* T,
                                                const double* C, const double* phi,
                                                double* partials) {
    double v[kRedMax] = {0, 0, 0, 0};
    uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const bool is_max = reduce_is_max(which);
    if (is_max)
        for (int k = 0; k < kRedMax; k++) v[k] = -1e300;
    if (t < a.N && a.flag[t] != kIsWalls) {
        const double zero3[3] = {0, 0, 0};
        reduce_value(a, which, t, rho ? rho[t] : 0.0, u ? u + 3 * (size_t)t : zero3,
                     T ? T[t] : 0.0,
                     C ? C[t] : 0.0, phi ? phi[t] : 0.0, [&](uint32_t k) { return phi[k]; }, v);
    }
    block_reduce(v, kRedMax, is_max, partials + (size_t)blockIdx.x * kRedMax);
}

__global__ void __launch_bounds__(256) k_reduce_final(const double* partials, uint32_t nb,
                                                      int is_max, double* out) {
    reduce_partials(partials, nb, is_max != 0, out);
}

template <class R>
__global__ void __launch_bounds__(256) k_nonfinite(const R* p, size_t n, double* partials) {
    nonfinite_partial<R>(p, n, partials);
}

int launch_reduce(int which, const StepArgs& a, const double* rho, const double* u,
                  const double* T, const double* C, const double* phi, double* partials,
                  double* out, int* nout, hipStream_t s) {
    int is_max = reduce_is_max(which) ? 1 : 0;
    k_reduce<<<grid_of(a), 256, 0, s>>>(a, which, rho, u, T, C, phi, partials);
    k_reduce_final<<<1, 256, 0, s>>>(partials, a.nblk, is_max, out);
    *nout = (which == 6 || which == 9) ? 3 : 1;
    return 0;
}

int launch_nonfinite(int fp32, const void* p, size_t n, double* partials, double* out,
                     hipStream_t s) {
    const uint32_t nb = 1024;
    if (fp32) k_nonfinite<float><<<nb, 256, 0, s>>>((const float*)p, n, partials);
    else k_nonfinite<double><<<nb, 256, 0, s>>>((const double*)p, n, partials);
    k_reduce_final<<<1, 256, 0, s>>>(partials, nb, 0, out);
    return 0;
}

}  // namespace mkg

// Host evaluation of the anisotropy of one crystal (mkPhaseField / mkMultiphaseField
// anisotropy, anisotropy_facet, anisotropy_cbc, anisotropy_hcp): the forms of the kernels.
extern "C" int mkg_aniso_eval(int kind, double eps4, const double* par, const double* rot,
                              const double* normal, double* a_s, double* n_v) {
    if (!normal || !a_s || !n_v) {
        mkg::set_error("mkg_aniso_eval: normal, a_s and n_v must not be null");
        return -1;
    }
    if (kind < MKG_ANISO_CUBIC || kind > MKG_ANISO_FACET) {
        mkg::set_error("mkg_aniso_eval: kind must be one of MKG_ANISO_*");
        return -1;
    }
    const double g2 = normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2];
    if (!(g2 > 1e-30)) {  // no interface direction: isotropic
        *a_s = 1.0;
        n_v[0] = n_v[1] = n_v[2] = 0.0;
        return 1;
    }
    const double gn = std::sqrt(g2);
    const double l[3] = {normal[0] / gn, normal[1] / gn, normal[2] / gn};
    double c[3];  // crystal frame: R^T l
    for (int k = 0; k < 3; k++)
        c[k] = rot ? rot[k] * l[0] + rot[3 + k] * l[1] + rot[6 + k] * l[2] : l[k];
    const mkg::Aniso<double> an{kind, eps4, par ? par[0] : 0.0, par ? par[1] : 0.0};
    double D[3];
    const double as = mkg::aniso_eval<double>(an, c[0], c[1], c[2], D[0], D[1], D[2]);
    const double nD = c[0] * D[0] + c[1] * D[1] + c[2] * D[2];
    double t[3];
    for (int k = 0; k < 3; k++) t[k] = as * (D[k] - c[k] * nD);
    for (int k = 0; k < 3; k++)
        n_v[k] = rot ? rot[3 * k] * t[0] + rot[3 * k + 1] * t[1] + rot[3 * k + 2] * t[2] : t[k];
    *a_s = as;
    return 0;
}

// Streaming rate of the device: a copy kernel (one four-float vector per lane, nontemporal
// loads and stores, one launch over the whole buffer), timed with HIP events on a stream of
// its own.  The reference against which a bandwidth-bound kernel's rate is read.
namespace {
typedef float copy_f4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_copy4(const copy_f4* __restrict__ src,
                                               copy_f4* __restrict__ dst, size_t n4) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n4) __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
}
}  // namespace

extern "C" int mkg_stream_copy_rate(size_t bytes, int reps, double* gbs) {
    if (!gbs || bytes < 16 || reps < 1) {
        mkg::set_error("mkg_stream_copy_rate: bytes >= 16, reps >= 1 and gbs required");
        return -1;
    }
    const size_t n4 = bytes / 16;
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, n4 * 16);
    if (e == hipSuccess) e = hipMalloc(&b, n4 * 16);
    if (e == hipSuccess) e = hipStreamCreate(&s);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, n4 * 16, s);
    float ms = 0;
    if (e == hipSuccess) {
        const dim3 grid((unsigned)((n4 + 255) / 256)), block(256);
        k_copy4<<<grid, block, 0, s>>>((const copy_f4*)a, (copy_f4*)b, n4);  // warm-up
        hipEventRecord(e0, s);
        for (int r = 0; r < reps; r++) k_copy4<<<grid, block, 0, s>>>((const copy_f4*)a, (copy_f4*)b, n4);
        hipEventRecord(e1, s);
        e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) {
        mkg::set_error(std::string("mkg_stream_copy_rate: ") + hipGetErrorString(e));
        return -1;
    }
    *gbs = 2.0 * (double)(n4 * 16) * reps / (ms * 1e-3) / 1e9;
    return 0;
}
