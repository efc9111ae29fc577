// Building blocks of the flow adjoint shared by the flow-only reverse step (fadj.hip) and the
// flow part of the coupled reverse step (adjoint.hip): moments that fix the w-orthogonal
// parts of a population vector (the parts of the generalised moment collision of
// flow_collide), the parts of one direction, and deterministic block sums.
#pragma once
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


// Reverse of the generalised moment collision with Guo forcing of one cell (the collision of
// flow_collide without partial bounce-back): f the pre-collision populations (full values), F
// the body force, l the adjoint of the post-collision populations, ub_ext an adjoint of the
// velocity u = (sum c f + F/2)/rho from outside the collision (the scalar lattices that are
// advected by it).  Writes the adjoint of f to lam_out[i * N + id] and returns dL/dF and, with
// RATES, dL/d(s_nu, s_bulk, s_odd, s_even):
//   f* = feq + A (f - feq) + C G with A = sum_k (1 - r_k) P_k, C = sum_k (1 - r_k/2) P_k over
//   the w-orthogonal parts P_k (so P_k^T = W^-1 P_k W); feq and G depend on f through
//   rho = sum f and u:
//   lambda_f = A^T l + J_feq^T (I - A)^T l + J_G^T C^T l,  dL/dF = J_G,F^T C^T l + ubar/(2 rho),
//   dL/dr_k = -l . P_k (f - feq + G/2).
template <int Q, class R, bool RATES>
__device__ inline void collide_reverse(const StepArgs& a, const R* f, const R (&F)[3], const R* l,
                                       const R (&ub_ext)[3], R* lam_out, uint32_t N, uint32_t id,
                                       R (&dF)[3], R (&drate)[4]) {
    R rho = 0, j[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < Q; i++) {
        rho += f[i];
        j[0] += cx(i) * f[i];
        j[1] += cy(i) * f[i];
        j[2] += cz(i) * f[i];
    }
    const R F0 = F[0], F1 = F[1], F2 = F[2];
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
    if constexpr (RATES) {
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
    }
    drate[0] = r4;
    drate[1] = r5;
    drate[2] = r6;
    drate[3] = r7;
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
    // A^T l waits in LDS (each thread reads back only its own slots) until rb and ubar are
    // known: a store, load and store again of lam_out would make every direction wait for its
    // own global store
    __shared__ R atl[Q][256];
    const uint32_t tid = threadIdx.x;
    R rb = 0, ub0 = ub_ext[0], ub1 = ub_ext[1], ub2 = ub_ext[2], q1 = 0, q2 = 0, q3 = 0;
#pragma unroll
    for (int i = 0; i < Q; i++) {
        R ATl, CTl;
        atl_ctl(i, ATl, CTl);
        atl[i][tid] = ATl;
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
    dF[0] = q1 + (R)0.5 * ub0 * ir;
    dF[1] = q2 + (R)0.5 * ub1 * ir;
    dF[2] = q3 + (R)0.5 * ub2 * ir;
#pragma unroll
    for (int i = 0; i < Q; i++)
        lam_out[(size_t)i * N + id] =
            atl[i][tid] + rb + (ub0 * (cx(i) - u0) + ub1 * (cy(i) - u1) + ub2 * (cz(i) - u2)) * ir;
}

}  // namespace mkg
