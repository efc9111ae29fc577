// Discrete adjoint of the flow-only step (mkFluidsLBM::propagate / relaxation and the MRT
// overrides, MaterKit/inc/mkFluidsLBM.h:47-51, mkFluidsMRT.h:42-46): the generalised moment
// collision of flow_collide (momentum; traceless second order at s_nu; trace at s_bulk; odd
// and even ghosts at s_odd / s_even -- BGK and TRT are its special cases) with Guo forcing,
// then streaming with periodic, bounce (any wall velocity) and outflow faces and wall cells.
// Loss L = 1/2 sum over non-wall cells |u_N - u*|^2, u = (sum c f + F/2)/rho.  The forward
// populations of every level come from a tape that the forward step kernel writes itself
// (its pre-collision populations, deviations in the storage type, slot i at tape + i*S); the
// loss reads the final level in canonical FP64.  The reverse runs in the forward's precision
// (float adjoints for FP32 storage) with FP64 block sums.
//
// Reverse of one step, per non-wall cell y (lambda = adjoint of the pre-collision populations
// of the next level, l = adjoint of the post-collision populations of this level):
//   l_i(y) = lambda_i(y + c_i)        when no face is crossed and y + c_i is not a wall
//          = lambda_inv(i)(y)          when y + c_i is a wall or lies behind a bounce face
//          + lambda_i(y)               when y - c_i lies behind an outflow face
//   f* = feq + A (f - feq) + C G with A = sum_k (1 - r_k) P_k, C = sum_k (1 - r_k/2) P_k over
//   the w-orthogonal parts P_k (so P_k^T = W^-1 P_k W); feq and G depend on f through
//   rho = sum f and u = (j + F/2)/rho:
//   lambda_f = A^T l + J_feq^T (I - A)^T l + J_G^T C^T l,  dL/dF += J_G,F^T C^T l + ubar/(2 rho),
//   dL/dr_k = -l . P_k (f - feq + G/2).
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "device.h"
#include "fadj.h"

namespace mkg {

// Loss and the adjoint of the final populations; partials 0 (loss) and 1..3 (dL/dF).
template <int Q, class R>
__global__ void __launch_bounds__(256) k_fadj_loss(const StepArgs a, const double* fN,
                                                   const double* ut, R* lam, double* part) {
    const uint32_t N = a.N;
    Cell c;
    const bool in = locate_clamped(a, c);
    const bool act = in && a.flag[c.id] != kIsWalls;
    double v[4] = {0, 0, 0, 0};
    if (in) {
        double f[Q], rho = 0, j[3] = {0, 0, 0};
#pragma unroll
        for (int i = 0; i < Q; i++) {
            f[i] = fN[(size_t)i * N + c.id];
            rho += f[i];
            j[0] += cx(i) * f[i];
            j[1] += cy(i) * f[i];
            j[2] += cz(i) * f[i];
        }
        double u[3], du[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            u[d] = (j[d] + 0.5 * a.force[d]) / rho;
            du[d] = act ? u[d] - ut[3 * (size_t)c.id + d] : 0.0;
            v[0] += 0.5 * du[d] * du[d];
            v[1 + d] = du[d] / (2.0 * rho);
        }
#pragma unroll
        for (int i = 0; i < Q; i++) {
            const double s = du[0] * (cx(i) - u[0]) + du[1] * (cy(i) - u[1]) + du[2] * (cz(i) - u[2]);
            lam[(size_t)i * N + c.id] = (R)(s / rho);
        }
    }
    fadj_block_add<4>(v, part, a.nblk);
}

// One reverse step: lam_in = adjoint of level n+1, tape = level n; partials 1..3 (dL/dF) and
// 4..7 (dL/d s_nu, s_bulk, s_odd, s_even) accumulate.
// Minimum waves per SIMD asked of the register allocator for the FP32 reverse step: two
// (197 VGPRs, no spills); three spill 76 bytes of scratch per lane.
#ifndef MKG_FADJ_WAVES_F32
#define MKG_FADJ_WAVES_F32 2
#endif

template <int Q, class R>
__global__ void __launch_bounds__(256, sizeof(R) == 4 ? MKG_FADJ_WAVES_F32 : 1)
    k_fadj_step(const StepArgs a, const R* __restrict__ tape, const R* lam_in, R* lam_out,
                double* part) {
    const uint32_t N = a.N;
    Cell c;
    const bool in = locate_clamped(a, c);
    const bool act = in && a.flag[c.id] != kIsWalls;
    double gsum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (act) {
        const uint32_t np = (uint32_t)a.f.npmask & c.bits;
        R l[Q];
#pragma unroll
        for (int i = 0; i < Q; i++) {
            const int X = cx(i), Y = cy(i), Z = cz(i);
            const uint32_t hp = np & faces_of(X, Y, Z);
            R v = 0;
            if (hp == 0u) {
                const uint32_t t = nbr(c, X, Y, Z);
                v = a.flag[t] == kIsWalls ? lam_in[(size_t)inv(i) * N + c.id]
                                          : lam_in[(size_t)i * N + t];
            } else if (pick6(__builtin_ctz(hp), a.f.bc) == kBcBounce) {
                v = lam_in[(size_t)inv(i) * N + c.id];
            }
            const uint32_t hb = np & faces_of(-X, -Y, -Z);
            if (hb != 0u && pick6(__builtin_ctz(hb), a.f.bc) == kBcOutflow)
                v += lam_in[(size_t)i * N + c.id];
            l[i] = v;
        }
        R f[Q];
#pragma unroll
        for (int i = 0; i < Q; i++)
            f[i] = (R)(wgt<Q>(i) * a.f.ref) + *slot_at(tape + (size_t)i * a.S, c.id);
        const R F[3] = {(R)a.force[0], (R)a.force[1], (R)a.force[2]};
        const R ub_ext[3] = {0, 0, 0};
        R dF[3], drate[4];
        collide_reverse<Q, R, true>(a, f, F, l, ub_ext, lam_out, N, c.id, dF, drate);
        for (int k = 0; k < 3; k++) gsum[1 + k] = (double)dF[k];
        for (int k = 0; k < 4; k++) gsum[4 + k] = (double)drate[k];
    } else if (in) {
#pragma unroll
        for (int i = 0; i < Q; i++) lam_out[(size_t)i * N + c.id] = (R)0;
    }
    fadj_block_add<8>(gsum, part, a.nblk);
}

// The reverse runs in the forward's precision (fp32: float adjoints and arithmetic), with
// the FP64 tape and FP64 block sums.
template <class R>
static int fadj_loss_t(int q, const StepArgs& a, const double* fN, const double* ut, void* lam,
                       double* part, hipStream_t s) {
    if (q == 19) k_fadj_loss<19, R><<<a.nblk, 256, 0, s>>>(a, fN, ut, (R*)lam, part);
    else if (q == 27) k_fadj_loss<27, R><<<a.nblk, 256, 0, s>>>(a, fN, ut, (R*)lam, part);
    else return -1;
    return 0;
}
template <class R>
static int fadj_step_t(int q, const StepArgs& a, const void* tape, const void* lam_in,
                       void* lam_out, double* part, hipStream_t s) {
    if (q == 19)
        k_fadj_step<19, R><<<a.nblk, 256, 0, s>>>(a, (const R*)tape, (const R*)lam_in,
                                                  (R*)lam_out, part);
    else if (q == 27)
        k_fadj_step<27, R><<<a.nblk, 256, 0, s>>>(a, (const R*)tape, (const R*)lam_in,
                                                  (R*)lam_out, part);
    else return -1;
    return 0;
}
int launch_fadj_loss(int q, int fp32, const StepArgs& a, const double* fN, const double* ut,
                     void* lam, double* part, hipStream_t s) {
    return fp32 ? fadj_loss_t<float>(q, a, fN, ut, lam, part, s)
                : fadj_loss_t<double>(q, a, fN, ut, lam, part, s);
}
int launch_fadj_step(int q, int fp32, const StepArgs& a, const void* tape, const void* lam_in,
                     void* lam_out, double* part, hipStream_t s) {
    return fp32 ? fadj_step_t<float>(q, a, tape, lam_in, lam_out, part, s)
                : fadj_step_t<double>(q, a, tape, lam_in, lam_out, part, s);
}

}  // namespace mkg
