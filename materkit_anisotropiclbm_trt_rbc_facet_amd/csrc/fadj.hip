// Discrete adjoint of the flow-only step (mkFluidsLBM::propagate / relaxation and the MRT
// overrides, MaterKit/inc/mkFluidsLBM.h:47-51, mkFluidsMRT.h:42-46): the generalised moment
// collision of flow_collide (momentum; traceless second order at s_nu; trace at s_bulk; odd
// and even ghosts at s_odd / s_even -- BGK and TRT are its special cases) with Guo forcing,
// then streaming with periodic, bounce (any wall velocity) and outflow faces and wall cells.
// Loss L = 1/2 sum over non-wall cells |u_N - u*|^2, u = (sum c f + F/2)/rho.  The forward
// populations come from a tape of canonical FP64 values; the reverse runs in the forward's
// precision (float adjoints for FP32 storage) with FP64 block sums.
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

namespace mkg {

// Moments of a Q-vector that fix its w-orthogonal parts: mass m0, momentum j, the traceless
// second moment T (xx, yy, zz, xy, xz, yz) and the trace moment tr = sum (c^2 - 3 cs^2) v.
template <class R>
struct Mom {
    R m0 = 0, j0 = 0, j1 = 0, j2 = 0, T0 = 0, T1 = 0, T2 = 0, T3 = 0, T4 = 0, T5 = 0, tr = 0;
    __device__ inline void add(int i, R x) {
        const R X = (R)cx(i), Y = (R)cy(i), Z = (R)cz(i);
        m0 += x;
        j0 += X * x;
        j1 += Y * x;
        j2 += Z * x;
        T0 += X * X * x;
        T1 += Y * Y * x;
        T2 += Z * Z * x;
        T3 += X * Y * x;
        T4 += X * Z * x;
        T5 += Y * Z * x;
    }
    // second moments -> traceless part and trace moment
    __device__ inline void finish() {
        const R trace = T0 + T1 + T2, t3 = trace / (R)3;
        T0 -= t3;
        T1 -= t3;
        T2 -= t3;
        tr = trace - m0;
    }
};
// Parts of direction i of the vector with moments m and component v: momentum, traceless
// second order, trace, and the remainder (ghost).
template <int Q, class R>
__device__ inline void part_of(int i, const Mom<R>& m, R v, R& mo, R& sh, R& bk, R& gh) {
    const R w = (R)wgt<Q>(i);
    const R X = (R)cx(i), Y = (R)cy(i), Z = (R)cz(i);
    mo = (R)3 * w * (X * m.j0 + Y * m.j1 + Z * m.j2);
    sh = (R)4.5 * w * (X * X * m.T0 + Y * Y * m.T1 + Z * Z * m.T2 + (R)2 * X * Y * m.T3 +
                       (R)2 * X * Z * m.T4 + (R)2 * Y * Z * m.T5);
    bk = (R)1.5 * w * (R)(csq(i) - 1) * m.tr;
    gh = v - w * m.m0 - mo - sh - bk;
}

// Deterministic block sum of NV values per thread (wave shuffles, then the four waves in
// order); thread 0 adds the totals to part[k * nblk + block].
template <int NV>
__device__ inline void fadj_block_add(double (&v)[NV], double* part, uint32_t nblk) {
    __shared__ double red[NV][4];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double x = v[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++)
            part[(size_t)k * nblk + blockIdx.x] += red[k][0] + red[k][1] + red[k][2] + red[k][3];
    }
}

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

// One reverse step: lam_in = adjoint of level n+1, fn = tape of level n; partials 1..3
// (dL/dF) and 4..7 (dL/d s_nu, s_bulk, s_odd, s_even) accumulate.
// Minimum waves per SIMD asked of the register allocator for the FP32 reverse step: two
// (240 VGPRs, no spills) measured faster than three or four, which spill (76 and 232 bytes
// of scratch per lane), at 128 x 64 x 64.
#ifndef MKG_FADJ_WAVES_F32
#define MKG_FADJ_WAVES_F32 2
#endif

template <int Q, class R>
__global__ void __launch_bounds__(256, sizeof(R) == 4 ? MKG_FADJ_WAVES_F32 : 1) k_fadj_step(const StepArgs a, const double* fn,
                                                   const R* lam_in, R* lam_out, double* part) {
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
        R f[Q], rho = 0, j[3] = {0, 0, 0};
#pragma unroll
        for (int i = 0; i < Q; i++) {
            f[i] = (R)fn[(size_t)i * N + c.id];
            rho += f[i];
            j[0] += cx(i) * f[i];
            j[1] += cy(i) * f[i];
            j[2] += cz(i) * f[i];
        }
        const R F0 = (R)a.force[0], F1 = (R)a.force[1], F2 = (R)a.force[2];
        const R ir = (R)1 / rho;
        const R u0 = (j[0] + (R)0.5 * F0) * ir, u1 = (j[1] + (R)0.5 * F1) * ir,
                u2 = (j[2] + (R)0.5 * F2) * ir;
        const R uu = u0 * u0 + u1 * u1 + u2 * u2, uF = u0 * F0 + u1 * F1 + u2 * F2;
        const R s_nu = (R)a.s_nu, s_b = (R)a.s_bulk, s_o = (R)a.s_odd, s_e = (R)a.s_even;
        auto feq_of = [&](int i) {
            const R cu = cx(i) * u0 + cy(i) * u1 + cz(i) * u2;
            return (R)wgt<Q>(i) * rho * ((R)1 + (R)3 * cu + (R)4.5 * cu * cu - (R)1.5 * uu);
        };
        // d = f - feq + G/2 (for the rates; f is dead afterwards)
        auto d_of = [&](int i) {
            const R cu = cx(i) * u0 + cy(i) * u1 + cz(i) * u2;
            const R cF = cx(i) * F0 + cy(i) * F1 + cz(i) * F2;
            return f[i] - feq_of(i) + (R)0.5 * (R)wgt<Q>(i) * ((R)3 * (cF - uF) + (R)9 * cu * cF);
        };
        Mom<R> md, ml;
#pragma unroll
        for (int i = 0; i < Q; i++) {
            md.add(i, d_of(i));
            ml.add(i, (R)wgt<Q>(i) * l[i]);
        }
        md.finish();
        ml.finish();
        // rates: dL/dr_k = -l . P_k (f - feq + G/2)
        R r4 = 0, r5 = 0, r6 = 0, r7 = 0;
#pragma unroll
        for (int i = 0; i < Q; i++) {
            R mo, sh, bk, gi, mk, sk, bkk, gk;
            part_of<Q, R>(i, md, d_of(i), mo, sh, bk, gi);
            part_of<Q, R>(inv(i), md, d_of(inv(i)), mk, sk, bkk, gk);
            r4 -= l[i] * sh;
            r5 -= l[i] * bk;
            r6 -= l[i] * (R)0.5 * (gi - gk);
            r7 -= l[i] * (R)0.5 * (gi + gk);
        }
        gsum[4] = (double)r4;
        gsum[5] = (double)r5;
        gsum[6] = (double)r6;
        gsum[7] = (double)r7;
        // A^T l and C^T l of direction i through P^T = W^-1 P W
        auto atl_ctl = [&](int i, R& ATl, R& CTl) {
            R mo, sh, bk, gi, mk, sk, bkk, gk;
            part_of<Q, R>(i, ml, (R)wgt<Q>(i) * l[i], mo, sh, bk, gi);
            part_of<Q, R>(inv(i), ml, (R)wgt<Q>(inv(i)) * l[inv(i)], mk, sk, bkk, gk);
            const R go = (R)0.5 * (gi - gk), ge = (R)0.5 * (gi + gk);
            const R iw = (R)(1.0 / wgt<Q>(i));
            ATl = (((R)1 - s_nu) * sh + ((R)1 - s_b) * bk + ((R)1 - s_o) * go + ((R)1 - s_e) * ge) * iw;
            CTl = ((R)0.5 * mo + ((R)1 - (R)0.5 * s_nu) * sh + ((R)1 - (R)0.5 * s_b) * bk +
                   ((R)1 - (R)0.5 * s_o) * go + ((R)1 - (R)0.5 * s_e) * ge) * iw;
        };
        R rb = 0, ub0 = 0, ub1 = 0, ub2 = 0, q1 = 0, q2 = 0, q3 = 0;
#pragma unroll
        for (int i = 0; i < Q; i++) {
            R ATl, CTl;
            atl_ctl(i, ATl, CTl);
            // the final adjoint's A^T l part now, the rest once rb and ubar are known
            lam_out[(size_t)i * N + c.id] = ATl;
            const R w = (R)wgt<Q>(i);
            const R X = (R)cx(i), Y = (R)cy(i), Z = (R)cz(i);
            const R cu = X * u0 + Y * u1 + Z * u2;
            const R cF = X * F0 + Y * F1 + Z * F2;
            const R mu = l[i] - ATl;
            rb += mu * feq_of(i) * ir;
            const R a1 = mu * w * rho, a2 = CTl * w;
            ub0 += a1 * ((R)3 * X + (R)9 * cu * X - (R)3 * u0) + a2 * ((R)-3 * F0 + (R)9 * X * cF);
            ub1 += a1 * ((R)3 * Y + (R)9 * cu * Y - (R)3 * u1) + a2 * ((R)-3 * F1 + (R)9 * Y * cF);
            ub2 += a1 * ((R)3 * Z + (R)9 * cu * Z - (R)3 * u2) + a2 * ((R)-3 * F2 + (R)9 * Z * cF);
            q1 += a2 * ((R)3 * (X - u0) + (R)9 * cu * X);
            q2 += a2 * ((R)3 * (Y - u1) + (R)9 * cu * Y);
            q3 += a2 * ((R)3 * (Z - u2) + (R)9 * cu * Z);
        }
        gsum[1] = (double)(q1 + (R)0.5 * ub0 * ir);
        gsum[2] = (double)(q2 + (R)0.5 * ub1 * ir);
        gsum[3] = (double)(q3 + (R)0.5 * ub2 * ir);
#pragma unroll
        for (int i = 0; i < Q; i++) {
            R* o = &lam_out[(size_t)i * N + c.id];
            *o = *o + rb + (ub0 * (cx(i) - u0) + ub1 * (cy(i) - u1) + ub2 * (cz(i) - u2)) * ir;
        }
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
static int fadj_step_t(int q, const StepArgs& a, const double* fn, const void* lam_in,
                       void* lam_out, double* part, hipStream_t s) {
    if (q == 19)
        k_fadj_step<19, R><<<a.nblk, 256, 0, s>>>(a, fn, (const R*)lam_in, (R*)lam_out, part);
    else if (q == 27)
        k_fadj_step<27, R><<<a.nblk, 256, 0, s>>>(a, fn, (const R*)lam_in, (R*)lam_out, part);
    else return -1;
    return 0;
}
int launch_fadj_loss(int q, int fp32, const StepArgs& a, const double* fN, const double* ut,
                     void* lam, double* part, hipStream_t s) {
    return fp32 ? fadj_loss_t<float>(q, a, fN, ut, lam, part, s)
                : fadj_loss_t<double>(q, a, fN, ut, lam, part, s);
}
int launch_fadj_step(int q, int fp32, const StepArgs& a, const double* fn, const void* lam_in,
                     void* lam_out, double* part, hipStream_t s) {
    return fp32 ? fadj_step_t<float>(q, a, fn, lam_in, lam_out, part, s)
                : fadj_step_t<double>(q, a, fn, lam_in, lam_out, part, s);
}

}  // namespace mkg
