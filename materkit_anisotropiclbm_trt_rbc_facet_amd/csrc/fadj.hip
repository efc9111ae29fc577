// Discrete adjoint of the flow-only step (mkFluidsLBM::propagate / relaxation and the MRT
// overrides, MaterKit/inc/mkFluidsLBM.h:47-51, mkFluidsMRT.h:42-46): the generalised moment
// collision of flow_collide (momentum; traceless second order at s_nu; trace at s_bulk; odd
// and even ghosts at s_odd / s_even -- BGK and TRT are its special cases) with Guo forcing,
// then streaming with periodic, bounce (any wall velocity) and outflow faces and wall cells.
// Loss L = 1/2 sum over non-wall cells |u_N - u*|^2, u = (sum c f + F/2)/rho.  The forward
// populations come from a tape of canonical FP64 values, and the reverse runs in FP64.
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

// Parts of a Q-vector v in the w-orthogonal decomposition: momentum, traceless second order,
// trace, odd and even ghosts (the mass part is implied: v minus all of them).
template <int Q>
__device__ inline void fadj_parts(const double (&v)[Q], double (&mo)[Q], double (&sh)[Q],
                                  double (&bk)[Q], double (&go)[Q], double (&ge)[Q]) {
    double m0 = 0, j[3] = {0, 0, 0}, P[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const double x = v[i];
        const double X = cx(i), Y = cy(i), Z = cz(i);
        m0 += x;
        j[0] += X * x;
        j[1] += Y * x;
        j[2] += Z * x;
        P[0] += X * X * x;
        P[1] += Y * Y * x;
        P[2] += Z * Z * x;
        P[3] += X * Y * x;
        P[4] += X * Z * x;
        P[5] += Y * Z * x;
    }
    const double trace = P[0] + P[1] + P[2];
    const double tr = trace - m0;  // sum (c^2 - 3 cs^2) v
    const double t3 = trace / 3.0;
    const double T0 = P[0] - t3, T1 = P[1] - t3, T2 = P[2] - t3;
    double g[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const double w = wgt<Q>(i);
        const double X = cx(i), Y = cy(i), Z = cz(i);
        mo[i] = 3.0 * w * (X * j[0] + Y * j[1] + Z * j[2]);
        sh[i] = 4.5 * w * (X * X * T0 + Y * Y * T1 + Z * Z * T2 + 2.0 * X * Y * P[3] +
                           2.0 * X * Z * P[4] + 2.0 * Y * Z * P[5]);
        bk[i] = 1.5 * w * (double)(csq(i) - 1) * tr;
        g[i] = v[i] - w * m0 - mo[i] - sh[i] - bk[i];
    }
#pragma unroll
    for (int i = 0; i < Q; i++) {
        go[i] = 0.5 * (g[i] - g[inv(i)]);
        ge[i] = 0.5 * (g[i] + g[inv(i)]);
    }
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
template <int Q>
__global__ void __launch_bounds__(256) k_fadj_loss(const StepArgs a, const double* fN,
                                                   const double* ut, double* lam, double* part) {
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
            lam[(size_t)i * N + c.id] = s / rho;
        }
    }
    fadj_block_add<4>(v, part, a.nblk);
}

// One reverse step: lam_in = adjoint of level n+1, fn = tape of level n; partials 1..3
// (dL/dF) and 4..7 (dL/d s_nu, s_bulk, s_odd, s_even) accumulate.
template <int Q>
__global__ void __launch_bounds__(256) k_fadj_step(const StepArgs a, const double* fn,
                                                   const double* lam_in, double* lam_out,
                                                   double* part) {
    const uint32_t N = a.N;
    Cell c;
    const bool in = locate_clamped(a, c);
    const bool act = in && a.flag[c.id] != kIsWalls;
    double gsum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (act) {
        const uint32_t np = (uint32_t)a.f.npmask & c.bits;
        double l[Q];
#pragma unroll
        for (int i = 0; i < Q; i++) {
            const int X = cx(i), Y = cy(i), Z = cz(i);
            const uint32_t hp = np & faces_of(X, Y, Z);
            double v = 0;
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
        double f[Q], rho = 0, j[3] = {0, 0, 0};
#pragma unroll
        for (int i = 0; i < Q; i++) {
            f[i] = fn[(size_t)i * N + c.id];
            rho += f[i];
            j[0] += cx(i) * f[i];
            j[1] += cy(i) * f[i];
            j[2] += cz(i) * f[i];
        }
        const double F0 = a.force[0], F1 = a.force[1], F2 = a.force[2];
        const double u0 = (j[0] + 0.5 * F0) / rho, u1 = (j[1] + 0.5 * F1) / rho,
                     u2 = (j[2] + 0.5 * F2) / rho;
        const double uu = u0 * u0 + u1 * u1 + u2 * u2, uF = u0 * F0 + u1 * F1 + u2 * F2;
        const double s_nu = a.s_nu, s_b = a.s_bulk, s_o = a.s_odd, s_e = a.s_even;
        double feq[Q], d[Q], lw[Q];
#pragma unroll
        for (int i = 0; i < Q; i++) {
            const double w = wgt<Q>(i);
            const double cu = cx(i) * u0 + cy(i) * u1 + cz(i) * u2;
            const double cF = cx(i) * F0 + cy(i) * F1 + cz(i) * F2;
            feq[i] = w * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
            const double G = w * (3.0 * (cF - uF) + 9.0 * cu * cF);
            d[i] = f[i] - feq[i] + 0.5 * G;
            lw[i] = w * l[i];
        }
        double mo[Q], sh[Q], bk[Q], go[Q], ge[Q];
        // rates: dL/dr_k = -l . P_k (f - feq + G/2)
        fadj_parts<Q>(d, mo, sh, bk, go, ge);
#pragma unroll
        for (int i = 0; i < Q; i++) {
            gsum[4] -= l[i] * sh[i];
            gsum[5] -= l[i] * bk[i];
            gsum[6] -= l[i] * go[i];
            gsum[7] -= l[i] * ge[i];
        }
        // A^T l and C^T l through P^T = W^-1 P W
        fadj_parts<Q>(lw, mo, sh, bk, go, ge);
        double rb = 0, ub0 = 0, ub1 = 0, ub2 = 0;
        double ATl[Q];
#pragma unroll
        for (int i = 0; i < Q; i++) {
            const double w = wgt<Q>(i);
            const double iw = 1.0 / w;
            ATl[i] = ((1.0 - s_nu) * sh[i] + (1.0 - s_b) * bk[i] + (1.0 - s_o) * go[i] +
                      (1.0 - s_e) * ge[i]) * iw;
            const double CTl = (0.5 * mo[i] + (1.0 - 0.5 * s_nu) * sh[i] + (1.0 - 0.5 * s_b) * bk[i] +
                                (1.0 - 0.5 * s_o) * go[i] + (1.0 - 0.5 * s_e) * ge[i]) * iw;
            const double mu = l[i] - ATl[i];
            const double X = cx(i), Y = cy(i), Z = cz(i);
            const double cu = X * u0 + Y * u1 + Z * u2;
            const double cF = X * F0 + Y * F1 + Z * F2;
            rb += mu * feq[i] / rho;
            const double a1 = mu * w * rho, a2 = CTl * w;
            ub0 += a1 * (3.0 * X + 9.0 * cu * X - 3.0 * u0) + a2 * (-3.0 * F0 + 9.0 * X * cF);
            ub1 += a1 * (3.0 * Y + 9.0 * cu * Y - 3.0 * u1) + a2 * (-3.0 * F1 + 9.0 * Y * cF);
            ub2 += a1 * (3.0 * Z + 9.0 * cu * Z - 3.0 * u2) + a2 * (-3.0 * F2 + 9.0 * Z * cF);
            gsum[1] += a2 * (3.0 * (X - u0) + 9.0 * cu * X);
            gsum[2] += a2 * (3.0 * (Y - u1) + 9.0 * cu * Y);
            gsum[3] += a2 * (3.0 * (Z - u2) + 9.0 * cu * Z);
        }
        gsum[1] += ub0 / (2.0 * rho);
        gsum[2] += ub1 / (2.0 * rho);
        gsum[3] += ub2 / (2.0 * rho);
#pragma unroll
        for (int i = 0; i < Q; i++)
            lam_out[(size_t)i * N + c.id] =
                ATl[i] + rb + (ub0 * (cx(i) - u0) + ub1 * (cy(i) - u1) + ub2 * (cz(i) - u2)) / rho;
    } else if (in) {
#pragma unroll
        for (int i = 0; i < Q; i++) lam_out[(size_t)i * N + c.id] = 0.0;
    }
    fadj_block_add<8>(gsum, part, a.nblk);
}

int launch_fadj_loss(int q, const StepArgs& a, const double* fN, const double* ut, double* lam,
                     double* part, hipStream_t s) {
    if (q == 19) k_fadj_loss<19><<<a.nblk, 256, 0, s>>>(a, fN, ut, lam, part);
    else if (q == 27) k_fadj_loss<27><<<a.nblk, 256, 0, s>>>(a, fN, ut, lam, part);
    else return -1;
    return 0;
}
int launch_fadj_step(int q, const StepArgs& a, const double* fn, const double* lam_in,
                     double* lam_out, double* part, hipStream_t s) {
    if (q == 19) k_fadj_step<19><<<a.nblk, 256, 0, s>>>(a, fn, lam_in, lam_out, part);
    else if (q == 27) k_fadj_step<27><<<a.nblk, 256, 0, s>>>(a, fn, lam_in, lam_out, part);
    else return -1;
    return 0;
}

}  // namespace mkg
