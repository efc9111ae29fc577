// Lattice constants shared by host and device code (D3Q7, D3Q19, D3Q27).
//
// Direction order: rest, the six faces (+x,-x,+y,-y,+z,-z), the twelve edges, the eight
// corners.  Opposite directions are the pairs (2k-1, 2k), so inv(i) is a bit flip away.
// The D3Q19 set is the first 19 entries of the D3Q27 set, D3Q7 the first 7.  The reference
// keeps the same tables in mkLattice (e, c, weight, inv; MaterKit/inc/mkLattice.h:33-50),
// with a separate scalar lattice (weight_scalar, cs_squ_scalar).
#pragma once
#include <hip/hip_runtime.h>

namespace mkg {

__host__ __device__ constexpr int cx(int i) {
    constexpr int t[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1,
                           -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
    return t[i];
}
__host__ __device__ constexpr int cy(int i) {
    constexpr int t[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0,
                           0, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
    return t[i];
}
__host__ __device__ constexpr int cz(int i) {
    constexpr int t[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1,
                           1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
    return t[i];
}
__host__ __device__ constexpr int inv(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }
__host__ __device__ constexpr int csq(int i) {
    return cx(i) * cx(i) + cy(i) * cy(i) + cz(i) * cz(i);
}

template <int Q>
__host__ __device__ constexpr double wgt(int i) {
    return Q == 7    ? (i == 0 ? 0.25 : 0.125)
           : Q == 19 ? (csq(i) == 0 ? 1.0 / 3.0 : (csq(i) == 1 ? 1.0 / 18.0 : 1.0 / 36.0))
                     : (csq(i) == 0   ? 8.0 / 27.0
                        : csq(i) == 1 ? 2.0 / 27.0
                        : csq(i) == 2 ? 1.0 / 54.0
                                      : 1.0 / 216.0);
}
template <int Q>
__host__ __device__ constexpr double cs2() {
    return Q == 7 ? 0.25 : 1.0 / 3.0;
}

// Node flags, identical codes to MaterKit/inc/mkGlobalDefs.h:140-145.
constexpr int kIsOutlet = -2, kIsInlet = -1, kIsFluid = 0, kIsInter = 1, kIsSolid = 2,
              kIsWalls = 3;

// Face boundary codes (faces x-, x+, y-, y+, z-, z+).
constexpr int kBcPeriodic = 0, kBcBounce = 1, kBcDirichlet = 2, kBcOutflow = 3, kBcPressure = 4,
              kBcVelocity = 5;

}  // namespace mkg
