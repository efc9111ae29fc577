// Per-cell macroscopic values and the FP64 block reduction shared by the reduction kernels
// (aux.hip, which reduces from per-cell double fields, and reduce_static.hip, which reduces
// straight from the population storage).  Both compute every cell value with the same
// functions, so their results agree bit for bit.  The diagnostics are those of
// Thermal_NusseltRBC / check_variables (MaterKit/inc/mkFluidsLBM.h:55,67-68).
#pragma once
#include "device.h"

namespace mkg {

constexpr int kRedMax = 4;

// Layout-aware population read: wall cells are never touched by a step, so their populations
// stay in their own slots.
template <class R, int Q>
__device__ inline void read_any(R* p, const LatArgs& L, const StepArgs& a, const Cell& c,
                                int kind, bool wall) {
    if (wall)
        read_pops<Q, false, R>(p, L, c, a, kind);
    else
        read_pops_rt<Q, R>(p, L, c, a, kind);
}

// Macroscopic values of one cell from its populations: T = sum h, C = sum c, rho = sum f and
// the Guo half-force velocity u = (sum c f + F/2) / rho (rho0 for the incompressible
// equilibrium), F the body force plus the Boussinesq term.  want_C: read the solute lattice.
struct CellMacro {
    double rho = 0, T = 0, C = 0;
    double u[3] = {0, 0, 0};
};

template <class R, int FQ>
__device__ inline CellMacro cell_macro(const StepArgs& a, const Cell& c, bool wall,
                                       bool want_C) {
    CellMacro m;
    if (a.h.pop) {
        R h[7];
        read_any<R, 7>(h, a.h, a, c, kKindScalar, wall);
        R s = 0;
#pragma unroll
        for (int i = 0; i < 7; i++) s += h[i];
        m.T = a.h.ref + (double)s;
    }
    if (a.c.pop && want_C) {
        R p[7];
        read_any<R, 7>(p, a.c, a, c, kKindScalar, wall);
        R s = 0;
#pragma unroll
        for (int i = 0; i < 7; i++) s += p[i];
        m.C = a.c.ref + (double)s;
    }
    if constexpr (FQ > 0) {
        R f[FQ];
        read_any<R, FQ>(f, a.f, a, c, kKindFlow, wall);
        R dr = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
        for (int i = 0; i < FQ; i++) {
            dr += f[i];
            if (cx(i)) jx += (R)cx(i) * f[i];
            if (cy(i)) jy += (R)cy(i) * f[i];
            if (cz(i)) jz += (R)cz(i) * f[i];
        }
        const R rh = (R)a.f.ref + dr;
        double F[3] = {a.force[0], a.force[1], a.force[2]};
        if (a.h.pop)
            for (int k = 0; k < 3; k++) F[k] += a.rho0 * a.buoy[k] * (m.T - a.T_ref);
        m.rho = (double)rh;
        const double den = a.incomp ? a.rho0 : (double)rh;
        m.u[0] = ((double)jx + 0.5 * F[0]) / den;
        m.u[1] = ((double)jy + 0.5 * F[1]) / den;
        m.u[2] = ((double)jz + 0.5 * F[2]) / den;
    }
    return m;
}

// Reductions with a maximum instead of a sum (MKG_R_UMAX, MKG_R_TIP).
__host__ __device__ inline bool reduce_is_max(int which) { return which == 7 || which == 9; }

// Value(s) of one non-wall cell for reduction `which` (the MKG_R_* codes).  phi_at(k) returns
// phi of the cell at linear index k (the +x / +y / +z neighbours for the tip).
template <class PhiAt>
__device__ inline void reduce_value(const StepArgs& a, int which, uint32_t t, double rho,
                                    const double* u, double T, double C, double phi,
                                    PhiAt phi_at, double* v) {
    switch (which) {
        case 0: v[0] = rho; break;
        case 1: v[0] = T; break;
        case 2: v[0] = C; break;
        case 3: {
            double fs = 0.5 * (1.0 + phi);
            v[0] = fs < 0 ? 0 : (fs > 1 ? 1 : fs);
            break;
        }
        case 4: v[0] = phi; break;
        case 5: v[0] = u[2] * T; break;
        case 6:
            for (int k = 0; k < 3; k++) v[k] = rho * u[k];
            break;
        case 7: v[0] = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]); break;
        case 8: v[0] = 0; break;
        case 9: {
            // phi = 0 crossings between a cell with phi >= 0 and its +x / +y / +z neighbour
            // with phi < 0, linearly interpolated; the largest per axis is the extreme point
            // of the phi = 0 surface along that axis (the tip)
            if (phi >= 0.0) {
                const uint32_t z = t % a.nz, r = t / a.nz, y = r % a.ny, x = r / a.ny;
                const uint32_t c3[3] = {x, y, z};
                const uint32_t n3[3] = {(uint32_t)a.nx, (uint32_t)a.ny, (uint32_t)a.nz};
                const uint32_t st[3] = {(uint32_t)a.ny * (uint32_t)a.nz, (uint32_t)a.nz, 1u};
                for (int k = 0; k < 3; k++) {
                    if (c3[k] + 1u < n3[k]) {
                        const double p1 = phi_at(t + st[k]);
                        if (p1 < 0.0) v[k] = c3[k] + phi / (phi - p1) - 0.5 * (n3[k] - 1.0);
                    }
                }
            }
            break;
        }
    }
}

// Wavefront (64-lane) butterfly, then LDS across the four wavefronts of a block; per-block
// partials are later combined by a single block in fixed order, so results are deterministic.
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ inline void block_reduce(double* v, int nv, bool is_max, double* out_block) {
    __shared__ double sm[4][kRedMax];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int k = 0; k < nv; k++) {
        double r = is_max ? wave_max(v[k]) : wave_sum(v[k]);
        if (lane == 0) sm[wv][k] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < nv; k++) {
            double r = sm[0][k];
            for (int w = 1; w < 4; w++) r = is_max ? fmax(r, sm[w][k]) : r + sm[w][k];
            out_block[k] = r;
        }
    }
}

// Combination of nb per-block partials by one block in a fixed order.
__device__ inline void reduce_partials(const double* partials, uint32_t nb, bool is_max,
                                       double* out) {
    double v[kRedMax];
    for (int k = 0; k < kRedMax; k++) v[k] = is_max ? -1e300 : 0.0;
    for (uint32_t b = threadIdx.x; b < nb; b += 256)
        for (int k = 0; k < kRedMax; k++) {
            double p = partials[(size_t)b * kRedMax + k];
            v[k] = is_max ? fmax(v[k], p) : v[k] + p;
        }
    block_reduce(v, kRedMax, is_max, out);
}

// Non-finite values of an n-element storage buffer, block partials (grid-stride).
template <class R>
__device__ inline void nonfinite_partial(const R* p, size_t n, double* partials) {
    double v[kRedMax] = {0, 0, 0, 0};
    size_t t = (size_t)blockIdx.x * 256u + threadIdx.x;
    size_t stride = (size_t)gridDim.x * 256u;
    for (size_t i = t; i < n; i += stride)
        if (!isfinite((double)p[i])) v[0] += 1.0;
    block_reduce(v, kRedMax, false, partials + (size_t)blockIdx.x * kRedMax);
}

}  // namespace mkg
