// Device-side building blocks of the time step: cell location, AA-pattern population access
// with face rules, and the per-cell collision / source functions.  Included by step.hip and
// aux.hip only.
//
// AA pattern (in-place streaming, one population array per lattice):
//   layout E (lay 0): own slot i of cell x holds the pre-collision f_i(x).  An E step reads
//     its own slots and writes post-collision f*_i to its own slot inv(i): purely local.
//   layout O (lay 1): f_i(x) sits in slot inv(i) of cell x - c_i.  An O step gathers from
//     there and scatters f*_i to slot i of cell x + c_i, which restores layout E.
// Every (cell, slot) is read and written by the same thread within a step, so the update
// needs no second array and no inter-thread ordering.  Domain faces and wall cells are only
// touched by O steps: a population whose source (read) or target (write) lies outside a
// non-periodic face or in a wall cell is redirected to the cell's own slots with the face
// rule (half-way bounce-back with the moving-wall term, anti-bounce-back for a Dirichlet
// value, or zero-gradient outflow).  The same rules in pull form are the streaming of
// mkFluidsUser::propagate (MaterKit/src/mkFluidsUser.cpp:77-97).
#pragma once
#include "kernels.h"
#include "lattice.h"

namespace mkg {

enum { kKindFlow = 0, kKindScalar = 1 };
enum : uint32_t { XLO = 1u, XHI = 2u, YLO = 4u, YHI = 8u, ZLO = 16u, ZHI = 32u };

struct Cell {
    uint32_t id;
    uint32_t ox[3], oy[3], oz[3];  // linear offsets of coordinate -1, 0, +1 (periodic wrap)
    uint32_t bits;                 // which domain faces the cell touches
    int x, y, z;
};

// Block order.  The hardware deals consecutive block ids round-robin to the eight XCDs, each
// with its own L2.  With xcd_chunk = K > 0 (and nblk a multiple of 8K), remap_block gives
// XCD x (block ids b with b mod 8 = x) runs of K consecutive logical blocks: its j-th block
// takes logical block ((j / K) 8 + x) K + j mod K.  Neighbour rows of a stencil then sit in the
// same XCD's L2 except at run ends, while the eight XCDs still sweep one window of the domain
// together.  K = nblk / 8 would give each XCD one contiguous eighth of the domain.
__device__ inline uint32_t remap_block(const StepArgs& a, uint32_t b, uint32_t nblk) {
    const uint32_t K = (uint32_t)a.xcd_chunk;
    if (K == 0u || nblk % (8u * K) != 0u) return b;
    const uint32_t x = b & 7u, j = b >> 3;
    return ((j / K) * 8u + x) * K + (j % K);
}

// Thread -> cell.  Blocks of 256 threads walk the linear index (z contiguous), in the block
// order of remap_block.
__device__ inline uint32_t thread_cell(const StepArgs& a) {
    return remap_block(a, blockIdx.x, a.nblk) * 256u + threadIdx.x;
}

__device__ inline void locate_at(const StepArgs& a, uint32_t t, Cell& c);

__device__ inline bool locate(const StepArgs& a, Cell& c) {
    const uint32_t t = thread_cell(a);
    if (t >= a.N) return false;
    locate_at(a, t, c);
    return true;
}

// For kernels whose threads must all reach block barriers: a thread past the last cell gets
// the last cell (so its loads stay in bounds) and false.
__device__ inline bool locate_clamped(const StepArgs& a, Cell& c) {
    const uint32_t t = thread_cell(a);
    locate_at(a, t < a.N ? t : a.N - 1u, c);
    return t < a.N;
}

__device__ inline void locate_at(const StepArgs& a, uint32_t t, Cell& c) {
    uint32_t nz = (uint32_t)a.nz, ny = (uint32_t)a.ny, nx = (uint32_t)a.nx;
    uint32_t z = t % nz;
    uint32_t r = t / nz;
    uint32_t y = r % ny;
    uint32_t x = r / ny;
    c.id = t;
    c.x = (int)x;
    c.y = (int)y;
    c.z = (int)z;
    uint32_t pl = ny * nz;
    c.ox[0] = (x == 0 ? nx - 1 : x - 1) * pl;
    c.ox[1] = x * pl;
    c.ox[2] = (x + 1 == nx ? 0 : x + 1) * pl;
    c.oy[0] = (y == 0 ? ny - 1 : y - 1) * nz;
    c.oy[1] = y * nz;
    c.oy[2] = (y + 1 == ny ? 0 : y + 1) * nz;
    c.oz[0] = (z == 0 ? nz - 1 : z - 1);
    c.oz[1] = z;
    c.oz[2] = (z + 1 == nz ? 0 : z + 1);
    c.bits = (x == 0 ? XLO : 0u) | (x + 1 == nx ? XHI : 0u) | (y == 0 ? YLO : 0u) |
             (y + 1 == ny ? YHI : 0u) | (z == 0 ? ZLO : 0u) | (z + 1 == nz ? ZHI : 0u);
}

constexpr uint32_t kAllFaces = 0x3Fu;

// Faces crossed when stepping from a cell by (dx, dy, dz).
__host__ __device__ constexpr uint32_t faces_of(int dx, int dy, int dz) {
    return (dx < 0 ? XLO : dx > 0 ? XHI : 0u) | (dy < 0 ? YLO : dy > 0 ? YHI : 0u) |
           (dz < 0 ? ZLO : dz > 0 ? ZHI : 0u);
}

// Select among six per-face values without indexing the kernel-argument block dynamically
// (a runtime index into the by-value argument struct would copy it to scratch memory).
template <class T>
__device__ inline T pick6(int f, const T* v) {
    T r = v[5];
    r = f == 4 ? v[4] : r;
    r = f == 3 ? v[3] : r;
    r = f == 2 ? v[2] : r;
    r = f == 1 ? v[1] : r;
    r = f == 0 ? v[0] : r;
    return r;
}
__device__ inline double face_cu(int f, int X, int Y, int Z, const double (*uf)[3]) {
    double r = 0;
#pragma unroll
    for (int k = 0; k < 6; k++)
        if (f == k) r = X * uf[k][0] + Y * uf[k][1] + Z * uf[k][2];
    return r;
}

// Rotating-wall part of the wall velocity of face f next to cell c: the tangential part of
// wall_omega x (r - wall_center), r the wall point half a cell through the face
// (bdyRotation / bdySpiral, MaterKit/inc/mkFluidsLBM.h:64-65).
template <class W>
__device__ inline void wall_rotation(const W& a, int f, const Cell& c, double* t) {
    const int ax = f >> 1;
    const double sg = (f & 1) ? 0.5 : -0.5;
    const double r0 = c.x + (ax == 0 ? sg : 0.0) - a.wall_center[0];
    const double r1 = c.y + (ax == 1 ? sg : 0.0) - a.wall_center[1];
    const double r2 = c.z + (ax == 2 ? sg : 0.0) - a.wall_center[2];
    const double* w = a.wall_omega;
    t[0] = ax == 0 ? 0.0 : w[1] * r2 - w[2] * r1;
    t[1] = ax == 1 ? 0.0 : w[2] * r0 - w[0] * r2;
    t[2] = ax == 2 ? 0.0 : w[0] * r1 - w[1] * r0;
}

// c_i . u_w for the wall of face f next to cell c: the face velocity plus the rotating part
// (W: StepArgs or FaceTab, which hold the same wall fields).
template <class W>
__device__ inline double wall_cu(const W& a, int f, const Cell& c, int X, int Y, int Z) {
    double cu = face_cu(f, X, Y, Z, a.u_face);
    if (a.wall_rot) {
        double t[3];
        wall_rotation(a, f, c, t);
        cu += X * t[0] + Y * t[1] + Z * t[2];
    }
    return cu;
}

__device__ inline uint32_t nbr(const Cell& c, int dx, int dy, int dz) {
    return c.ox[dx + 1] + c.oy[dy + 1] + c.oz[dz + 1];
}

// Neighbourhood of a cell for the gradient stencils: the offsets of coordinates -1, 0, +1
// along each axis, with the cell's own offset in place of the neighbour across each face in
// np (the edge value replicated at non-periodic faces).  Held as named scalars: arrays copied
// from the Cell record and patched at run time made the compiler keep the record in scratch
// memory.
struct Nbh {
    uint32_t x0, x1, x2, y0, y1, y2, z0, z1, z2;
    __device__ uint32_t at(int dx, int dy, int dz) const {
        return (dx < 0 ? x0 : (dx == 0 ? x1 : x2)) + (dy < 0 ? y0 : (dy == 0 ? y1 : y2)) +
               (dz < 0 ? z0 : (dz == 0 ? z1 : z2));
    }
};
__device__ inline Nbh replicated_nbh(const Cell& c, uint32_t np) {
    Nbh o;
    o.x1 = c.ox[1];
    o.y1 = c.oy[1];
    o.z1 = c.oz[1];
    o.x0 = (np & XLO) ? o.x1 : c.ox[0];
    o.x2 = (np & XHI) ? o.x1 : c.ox[2];
    o.y0 = (np & YLO) ? o.y1 : c.oy[0];
    o.y2 = (np & YHI) ? o.y1 : c.oy[2];
    o.z0 = (np & ZLO) ? o.z1 : c.oz[0];
    o.z2 = (np & ZHI) ? o.z1 : c.oz[2];
    return o;
}

// ------------------------------------------------------------------ population access
// Populations are touched exactly once per step, so their loads and stores bypass cache
// retention (nontemporal hints) and leave L2 to the phase stencil and the shifted lines of O
// steps; measured +7 % on the fused 256^3 step over plain accesses.
#ifndef MKG_NT_POPS
#define MKG_NT_POPS 1
#endif
#ifndef MKG_NT_LOAD
#define MKG_NT_LOAD MKG_NT_POPS
#endif
#ifndef MKG_NT_STORE
#define MKG_NT_STORE MKG_NT_POPS
#endif
template <class R>
__device__ inline R ldp(const R* p) {
#if MKG_NT_LOAD
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}
template <class R>
__device__ inline void stp(R* p, R v) {
#if MKG_NT_STORE
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

// Element idx of slot array `slot`.  The slot base is uniform (scalar registers) and the byte
// offset a 32-bit vector value, which global loads and stores take as their address directly,
// with no 64-bit address arithmetic per access.  The host keeps every slot array below 4 GiB
// (cell count times element size < 2^32).
template <class R>
__device__ inline R* slot_at(R* slot, uint32_t idx) {
    return (R*)((char*)slot + idx * (uint32_t)sizeof(R));
}
template <class R>
__device__ inline const R* slot_at(const R* slot, uint32_t idx) {
    return (const R*)((const char*)slot + idx * (uint32_t)sizeof(R));
}

// Population addressing (StepArgs, LatArgs): the element offset of cell c within a slot of
// lattice L, and slot i of a lattice.  In the slot-major layout pcell(c) = c.  FP32 pages hold
// every lattice (one stride for all, uniform across the lattices of a kernel); FP64 pages hold
// one lattice each.
template <class R>
__device__ inline uint32_t pcell(const StepArgs& a, const LatArgs& L, uint32_t c) {
    return (c >> a.pg_shift) * (sizeof(R) == 4 ? a.pg_stride : L.pg_stride) + (c & a.pg_mask);
}
template <class R>
__device__ inline R* pslot(R* base, int i, const StepArgs& a) {
    return base + (size_t)i * a.sl_stride;
}
template <class R>
__device__ inline const R* pslot(const R* base, int i, const StepArgs& a) {
    return base + (size_t)i * a.sl_stride;
}
// Population i of cell c of lattice L (base: L's slot 0).
template <class R>
__device__ inline R* pop_at(R* base, int i, uint32_t c, const StepArgs& a, const LatArgs& L) {
    return slot_at(pslot(base, i, a), pcell<R>(a, L, c));
}
template <class R>
__device__ inline const R* pop_at(const R* base, int i, uint32_t c, const StepArgs& a,
                                  const LatArgs& L) {
    return slot_at(pslot(base, i, a), pcell<R>(a, L, c));
}
// Face-buffer value of slot i for face cell k.
template <class R>
__device__ inline R* face_at(const LatArgs& L, int i, uint32_t F, uint32_t k) {
    return (R*)L.face + (size_t)i * F + k;
}

// Position of a face cell in the face buffer: faces x-, x+ (ny*nz cells each, index
// y*nz + z), then y-, y+ (x*nz + z), then z-, z+ (x*ny + y).
__device__ inline uint32_t face_slot(const StepArgs& a, const Cell& c, int f) {
    const uint32_t ny = (uint32_t)a.ny, nz = (uint32_t)a.nz, nx = (uint32_t)a.nx;
    const uint32_t A = ny * nz, B = nx * nz, C = nx * ny;
    const uint32_t side = (uint32_t)f & 1u;
    if (f < 2) return side * A + (uint32_t)c.y * nz + (uint32_t)c.z;
    if (f < 4) return 2u * A + side * B + (uint32_t)c.x * nz + (uint32_t)c.z;
    return 2u * (A + B) + side * C + (uint32_t)c.x * ny + (uint32_t)c.y;
}

// Cell state that enters the face rules: the liquid fraction 1 - n_s scales the inflow
// (moving-wall) term of the flow, so a solid cell on an inlet receives no mass; the Dirichlet
// value of a scalar is multiplied by dscale (A(phi) for the solute, whose faces fix U).
template <class R>
struct FaceState {
    R wscale = (R)1;
    R dscale = (R)1;
    R ux = 0, uy = 0, uz = 0;  // the cell's fluid velocity (pressure faces)
};

// Population entering the cell in direction i through face f, from the cell's post-collision
// populations p (the pull rule of mkFluidsUser::propagate, MaterKit/src/mkFluidsUser.cpp:77-97,
// with the face codes of the case):
//   outflow    zero gradient: the cell's own post-collision population i
//   Dirichlet  anti-bounce-back to the face value (scalar lattices)
//   bounce     half-way bounce-back, plus 6 w rho0 c.u_w for a moving wall or inlet (flow)
//   pressure   anti-bounce-back to the second-order equilibrium at rho0 and the cell's
//              velocity u: 2 w rho0 (1 + 4.5 (c.u)^2 - 1.5 u.u) - f*_inv(i) (flow outlet)
// The parameters come from the boundary table (a.ftab, entry L.fid), loaded where a face rule
// runs.
// The parameters of one lattice's rule on one face, taken once for all the directions that
// cross the face: the face code, the table values, the face velocity and the rotating part of
// the wall velocity at the cell.
struct FaceRule {
    int t;
    bool rotating;
    double ref, fval, rho0;
    double u[3], rot[3];
};
__device__ inline FaceRule face_rule(const StepArgs& a, const LatArgs& L, int f, const Cell& c) {
    const FaceTab& T = *a.ftab;
    FaceRule r;
    r.t = T.bc[L.fid][f];
    r.ref = T.ref[L.fid];
    r.fval = T.fval[L.fid][f];
    r.rho0 = T.rho0;
#pragma unroll
    for (int k = 0; k < 3; k++) r.u[k] = T.u_face[f][k];
    r.rotating = T.wall_rot != 0;
    if (r.rotating)
        wall_rotation(T, f, c, r.rot);
    else
        r.rot[0] = r.rot[1] = r.rot[2] = 0.0;
    return r;
}

template <int Q, class R>
__device__ inline R face_in(const R* p, int i, const FaceRule& r, int kind,
                            const FaceState<R>& fs) {
    // outflow, and the layer of a velocity face, which continues itself
    if (r.t == kBcOutflow || (r.t == kBcVelocity && kind == kKindFlow)) return p[i];
    if (r.t == kBcPressure) {
        const R cu = cx(i) * fs.ux + cy(i) * fs.uy + cz(i) * fs.uz;
        const R uu = fs.ux * fs.ux + fs.uy * fs.uy + fs.uz * fs.uz;
        const R e = (R)r.rho0 * ((R)1 + (R)4.5 * cu * cu - (R)1.5 * uu);
        return -p[inv(i)] + (R)(2.0 * wgt<Q>(i)) * (e - (R)r.ref);
    }
    if (r.t == kBcDirichlet && kind == kKindScalar)
        return -p[inv(i)] + (R)(2.0 * wgt<Q>(i)) * (fs.dscale * (R)r.fval - (R)r.ref);
    R v = p[inv(i)];
    if (kind == kKindFlow) {
        double cu = cx(i) * r.u[0] + cy(i) * r.u[1] + cz(i) * r.u[2];
        if (r.rotating) cu += cx(i) * r.rot[0] + cy(i) * r.rot[1] + cz(i) * r.rot[2];
        v += (R)(6.0 * wgt<Q>(i) * r.rho0 * cu) * fs.wscale;
    }
    return v;
}

template <int Q, class R>
__device__ inline R face_in(const R* p, int i, int f, const LatArgs& L, const StepArgs& a,
                            int kind, const FaceState<R>& fs, const Cell& c) {
    return face_in<Q, R>(p, i, face_rule(a, L, f, c), kind, fs);
}

// AA access.  E step (ODD false): read own slot i, write own slot inv(i).  O step: read slot
// inv(i) of x - c_i, write slot i of x + c_i.  A population that crosses a non-periodic face
// is settled by face_in with the state of the step that collided it:
//   - crossing out in an O step, it returns at once into the cell's own slot inv(i) (no other
//     thread writes that slot);
//   - crossing out in an E step, the E step stores the returning population in the face
//     buffer, which the face cell alone reads in the next O step.  Its own slots cannot hold
//     it: the interior neighbours overwrite them during the O step.
// FACES: the cell may lie on a non-periodic face or next to a wall cell.  FM (with FACES)
// limits the face checks to the faces in the mask at compile time: a block whose only face
// cells lie on z faces, in a domain without wall cells, takes FM = ZLO | ZHI and carries no
// x / y face or wall logic.
template <int Q, bool ODD, class R, bool FACES = true, uint32_t FM = kAllFaces>
__device__ inline void read_pops(R* p, const LatArgs& L, const Cell& c,
                                 const StepArgs& a, int kind) {
    constexpr bool WALLS = FACES && FM == kAllFaces;
    const R* __restrict__ base = (const R*)L.pop;
    if constexpr (!ODD) {
#pragma unroll
        for (int i = 0; i < Q; i++) p[i] = ldp(pop_at(base, i, c.id, a, L));
    } else {
        (void)kind;
#pragma unroll
        for (int i = 0; i < Q; i++) {
            const int dx = -cx(i), dy = -cy(i), dz = -cz(i);
            uint32_t hit = FACES ? c.bits & faces_of(dx, dy, dz) & FM & (uint32_t)L.npmask : 0u;
            if (hit == 0u) {
                uint32_t src = nbr(c, dx, dy, dz);
                if (WALLS && a.has_walls && i != 0 && a.flag[src] == kIsWalls)
                    p[i] = ldp(pop_at(base, i, c.id, a, L));
                else
                    p[i] = ldp(pop_at(base, inv(i), src, a, L));
            } else {
                const int face = __builtin_ctz(hit);
                p[i] = ldp(face_at<R>(L, i, a.F, face_slot(a, c, face)));
            }
        }
    }
}

template <int Q, bool ODD, class R, bool FACES = true, uint32_t FM = kAllFaces>
__device__ inline void write_pops(const R* p, const LatArgs& L, const Cell& c,
                                  const StepArgs& a, int kind,
                                  const FaceState<R>& fs = FaceState<R>()) {
    constexpr bool WALLS = FACES && FM == kAllFaces;
    R* __restrict__ base = (R*)L.pop;
    if constexpr (!ODD) {
#pragma unroll
        for (int i = 0; i < Q; i++) stp(pop_at(base, inv(i), c.id, a, L), p[i]);
        if constexpr (FACES && FM == (ZLO | ZHI)) {
            // z faces alone: the lane at z = 0 (z = nz - 1) stores the populations that will
            // enter through a non-periodic z- (z+) face, one branch per face with the face
            // known at compile time
            const uint32_t npz = (uint32_t)L.npmask & (ZLO | ZHI);
            if ((npz & ZLO) && c.z == 0) {
                const uint32_t k = face_slot(a, c, 4);
                const FaceRule fr = face_rule(a, L, 4, c);
#pragma unroll
                for (int i = 1; i < Q; i++)
                    if (cz(i) > 0) stp(face_at<R>(L, i, a.F, k), face_in<Q, R>(p, i, fr, kind, fs));
            }
            if ((npz & ZHI) && c.z == a.nz - 1) {
                const uint32_t k = face_slot(a, c, 5);
                const FaceRule fr = face_rule(a, L, 5, c);
#pragma unroll
                for (int i = 1; i < Q; i++)
                    if (cz(i) < 0) stp(face_at<R>(L, i, a.F, k), face_in<Q, R>(p, i, fr, kind, fs));
            }
        } else if (FACES && (c.bits & FM & (uint32_t)L.npmask)) {
#pragma unroll
            for (int i = 1; i < Q; i++) {
                const uint32_t hit =
                    c.bits & faces_of(-cx(i), -cy(i), -cz(i)) & FM & (uint32_t)L.npmask;
                if (hit) {
                    const int f = __builtin_ctz(hit);
                    stp(face_at<R>(L, i, a.F, face_slot(a, c, f)),
                        face_in<Q, R>(p, i, f, L, a, kind, fs, c));
                }
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < Q; i++) {
            const int dx = cx(i), dy = cy(i), dz = cz(i);
            uint32_t hit = FACES ? c.bits & faces_of(dx, dy, dz) & FM & (uint32_t)L.npmask : 0u;
            if (hit == 0u) {
                uint32_t tgt = nbr(c, dx, dy, dz);
                if (WALLS && a.has_walls && i != 0 && a.flag[tgt] == kIsWalls)
                    stp(pop_at(base, inv(i), c.id, a, L), p[i]);
                else
                    stp(pop_at(base, i, tgt, a, L), p[i]);
            } else {
                const int face = __builtin_ctz(hit);
                stp(pop_at(base, inv(i), c.id, a, L),
                    face_in<Q, R>(p, inv(i), face, L, a, kind, fs, c));
            }
        }
    }
}

// ---- O step with the shift along z (the contiguous axis) taken through LDS
// A population with c_z != 0 moves by one element along z.  Read or written directly, every
// wave access is misaligned by one element and the step runs at 0.87 (FP32) / 0.67 (FP64) of
// the E step's bandwidth.  When a 256-thread block holds whole z rows (256 % nz == 0), the O
// step instead reads and writes each such slot at the cell's own z (aligned) and moves the
// values by one row position through LDS, wrapping at the row end (periodic z; a non-periodic
// z face is settled by the face rule as before).

// index of direction i among the directions with c_z != 0
template <int Q>
__host__ __device__ constexpr int zslot(int i) {
    int k = 0;
    for (int j = 0; j < i; j++)
        if (cz(j) != 0) k++;
    return k;
}
template <int Q>
__host__ __device__ constexpr int nzdirs() {
    return zslot<Q>(Q);
}

// The pull and the push of an O step come in two halves each, separated by a block barrier
// that the caller places, so that several lattices share one barrier:
//   zpull_issue   issues every global load of the lattice: the aligned loads that feed the
//                 stage (stored to zv), the direct ones (c_z == 0) and the face / wall values;
//                 returns the mask of z directions whose value does not come through the stage
//   zpull_finish  (after a barrier) takes the shifted values from zv, wrapping at the row end
//   zpush_stage   stores the c_z == 0 populations and the face / wall returns directly, and
//                 puts the z populations and a per-thread "sent normally" mask in LDS
//   zpush_finish  (after a barrier) the thread of the target's row position (x+cx, y+cy, z)
//                 stores each z population aligned, finding it in the sender's stage slot
// Every thread of the block runs all four (act = the thread holds a live, non-wall cell).  A
// stage region must not be written again before every thread has passed a barrier that
// follows its last read.  Global memory needs no ordering: each (cell, slot) is loaded and
// stored by one thread.
template <int Q, class R, bool FACES = true, uint32_t FM = kAllFaces>
__device__ inline uint32_t zpull_issue(R* p, const LatArgs& L, const Cell& c,
                                       const StepArgs& a, R (*zv)[256], uint32_t tid) {
    constexpr bool WALLS = FACES && FM == kAllFaces;
    constexpr int NZ = nzdirs<Q>();
    static_assert(NZ <= 32, "z mask too small");
    const R* __restrict__ base = (const R*)L.pop;
    R t[NZ];
#pragma unroll
    for (int i = 1; i < Q; i++)
        if (cz(i) != 0)
            t[zslot<Q>(i)] = ldp(pop_at(base, inv(i), nbr(c, -cx(i), -cy(i), 0), a, L));
    uint32_t own = 0;
    if constexpr (FACES && FM == (ZLO | ZHI)) {
        // z faces alone: every lane takes the face-free pull; the lane at z = 0 (z = nz - 1)
        // then takes the populations arriving through a non-periodic z- (z+) face from the
        // face buffer, one branch per face with the face known at compile time
#pragma unroll
        for (int i = 0; i < Q; i++)
            if (cz(i) == 0) p[i] = ldp(pop_at(base, inv(i), nbr(c, -cx(i), -cy(i), 0), a, L));
        const uint32_t npz = (uint32_t)L.npmask & (ZLO | ZHI);
        if ((npz & ZLO) && c.z == 0) {
            const uint32_t k = face_slot(a, c, 4);
#pragma unroll
            for (int i = 1; i < Q; i++)
                if (cz(i) > 0) {
                    p[i] = ldp(face_at<R>(L, i, a.F, k));
                    own |= 1u << zslot<Q>(i);
                }
        }
        if ((npz & ZHI) && c.z == a.nz - 1) {
            const uint32_t k = face_slot(a, c, 5);
#pragma unroll
            for (int i = 1; i < Q; i++)
                if (cz(i) < 0) {
                    p[i] = ldp(face_at<R>(L, i, a.F, k));
                    own |= 1u << zslot<Q>(i);
                }
        }
#pragma unroll
        for (int k = 0; k < NZ; k++) zv[k][tid] = t[k];
        return own;
    }
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const int dx = -cx(i), dy = -cy(i), dz = -cz(i);
        const uint32_t hit =
            FACES ? c.bits & faces_of(dx, dy, dz) & FM & (uint32_t)L.npmask : 0u;
        if (hit) {
            p[i] = ldp(face_at<R>(L, i, a.F, face_slot(a, c, __builtin_ctz(hit))));
            if (dz != 0) own |= 1u << zslot<Q>(i);
        } else {
            const uint32_t src = nbr(c, dx, dy, dz);
            if (WALLS && a.has_walls && i != 0 && a.flag[src] == kIsWalls) {
                p[i] = ldp(pop_at(base, i, c.id, a, L));
                if (dz != 0) own |= 1u << zslot<Q>(i);
            } else if (dz == 0) {
                p[i] = ldp(pop_at(base, inv(i), src, a, L));
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NZ; k++) zv[k][tid] = t[k];
    return own;
}

template <int Q, class R>
__device__ inline void zpull_finish(R* p, const Cell& c, const StepArgs& a, R (*zv)[256],
                                    uint32_t tid, uint32_t own) {
    const uint32_t zm1 = (uint32_t)a.nz - 1u;
    const uint32_t rb = tid - (tid & zm1);  // first thread of this z row
#pragma unroll
    for (int i = 1; i < Q; i++)
        if (cz(i) != 0 && !((own >> zslot<Q>(i)) & 1u))
            p[i] = zv[zslot<Q>(i)][rb + ((uint32_t)(c.z - cz(i)) & zm1)];
}

template <int Q, class R, bool FACES = true, uint32_t FM = kAllFaces>
__device__ inline void zpush_stage(const R* p, const LatArgs& L, const Cell& c,
                                   const StepArgs& a, int kind, R (*zv)[256], uint32_t* zm,
                                   uint32_t tid, bool act, const FaceState<R>& fs) {
    constexpr bool WALLS = FACES && FM == kAllFaces;
    static_assert(nzdirs<Q>() <= 32, "z mask too small");
    R* __restrict__ base = (R*)L.pop;
    if constexpr (FACES && FM == (ZLO | ZHI)) {
        // z faces alone: every lane stages its z populations and stores the others directly;
        // the lane at z = 0 (z = nz - 1) then returns the populations that leave through a
        // non-periodic z- (z+) face into its own slots.  zpush_finish skips their wrap.
#pragma unroll
        for (int i = 0; i < Q; i++) {
            if (cz(i) != 0)
                zv[zslot<Q>(i)][tid] = p[i];
            else if (act)
                stp(pop_at(base, i, nbr(c, cx(i), cy(i), 0), a, L), p[i]);
        }
        const uint32_t npz = (uint32_t)L.npmask & (ZLO | ZHI);
        if ((npz & ZLO) && c.z == 0 && act) {
            const FaceRule fr = face_rule(a, L, 4, c);
#pragma unroll
            for (int i = 1; i < Q; i++)
                if (cz(i) < 0)
                    stp(pop_at(base, inv(i), c.id, a, L), face_in<Q, R>(p, inv(i), fr, kind, fs));
        }
        if ((npz & ZHI) && c.z == a.nz - 1 && act) {
            const FaceRule fr = face_rule(a, L, 5, c);
#pragma unroll
            for (int i = 1; i < Q; i++)
                if (cz(i) > 0)
                    stp(pop_at(base, inv(i), c.id, a, L), face_in<Q, R>(p, inv(i), fr, kind, fs));
        }
        return;
    }
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const int dx = cx(i), dy = cy(i), dz = cz(i);
        const uint32_t hit =
            FACES ? c.bits & faces_of(dx, dy, dz) & FM & (uint32_t)L.npmask : 0u;
        if (dz != 0) zv[zslot<Q>(i)][tid] = p[i];
        if (hit) {
            if (act)
                stp(pop_at(base, inv(i), c.id, a, L),
                    face_in<Q, R>(p, inv(i), __builtin_ctz(hit), L, a, kind, fs, c));
        } else {
            const uint32_t tgt = nbr(c, dx, dy, dz);
            if (WALLS && a.has_walls && i != 0 && a.flag[tgt] == kIsWalls) {
                if (act) stp(pop_at(base, inv(i), c.id, a, L), p[i]);
            } else if (dz == 0) {
                if (act) stp(pop_at(base, i, tgt, a, L), p[i]);
            } else {
                m |= 1u << zslot<Q>(i);
            }
        }
    }
    // without faces and walls, with every thread updating, every z population is sent
    if constexpr (FACES) zm[tid] = act ? m : 0u;
}

// FACES false: the block has no face cells, no walls and no thread past the last cell, so
// every sender sent every z population and the masks are not read.
// FM = ZLO | ZHI (no wall cells, no x / y faces): a row's cells are live or not together, and
// the only z population not sent is the one that crosses a non-periodic z face: the receiver
// at z = 0 takes no c_z = +1 value (its sender, z = nz - 1, met the z+ face) and the receiver
// at z = nz - 1 no c_z = -1 value (z- face); act is the receiver's own liveness.
template <int Q, class R, bool FACES = true, uint32_t FM = kAllFaces>
__device__ inline void zpush_finish(const LatArgs& L, const Cell& c,
                                    const StepArgs& a, R (*zv)[256], const uint32_t* zm,
                                    uint32_t tid, bool act = true) {
    constexpr bool ZROWS = FACES && FM == (ZLO | ZHI);
    R* __restrict__ base = (R*)L.pop;
    const uint32_t zm1 = (uint32_t)a.nz - 1u;
    const uint32_t rb = tid - (tid & zm1);
#pragma unroll
    for (int i = 1; i < Q; i++) {
        if (cz(i) != 0) {
            // sender: the cell (x, y, z - c_z) of this thread's row
            const uint32_t s = rb + ((uint32_t)(c.z - cz(i)) & zm1);
            bool send = true;
            if constexpr (ZROWS) {
                const uint32_t zr = cz(i) > 0 ? 0u : zm1;
                const uint32_t face = cz(i) > 0 ? (uint32_t)ZHI : (uint32_t)ZLO;
                send = act && !((uint32_t)c.z == zr && ((uint32_t)L.npmask & face));
            } else if constexpr (FACES) {
                send = (zm[s] >> zslot<Q>(i)) & 1u;
            }
            if (send) stp(pop_at(base, i, nbr(c, cx(i), cy(i), 0), a, L), zv[zslot<Q>(i)][s]);
        }
    }
}

// Read with a runtime layout (lattices used only for coupling in this launch).
template <int Q, class R>
__device__ inline void read_pops_rt(R* p, const LatArgs& L, const Cell& c,
                                    const StepArgs& a, int kind) {
    if (L.lay)
        read_pops<Q, true, R>(p, L, c, a, kind);
    else
        read_pops<Q, false, R>(p, L, c, a, kind);
}

// ------------------------------------------------------------------ flow collision
// Generalised moment relaxation on the w-orthogonal Hermite decomposition of f - f_eq:
// momentum part (rate irrelevant under Guo forcing), traceless second-order "shear" part
// (s_nu), trace "bulk" part (s_bulk) and the remaining ghost moments split into odd (s_odd)
// and even (s_even) halves.  BGK is all four rates equal; TRT is s_bulk = s_even = s_nu,
// s_odd = omega-; MRT uses independent rates.  Forcing is Guo et al. (2002) projected onto
// the same parts with (1 - s/2) factors.  Populations are deviations f - w_i*ref; dr is their
// sum (rho - ref); rho is the density of the velocity terms of the equilibrium (rho itself, or
// rho0 for the He-Luo incompressible form).  Counterparts: mkNode::collision_BGK / collision_TRT / collision_MRT0-2
// (MaterKit/inc/mkNode.h:72-87).
template <int Q, class R>
__device__ inline void flow_collide(R* f, R dr, R rho, R ux, R uy, R uz, R Fx, R Fy, R Fz,
                                    R s_nu, R s_b, R s_o, R s_e, R ns) {
    // Second moments of the stored deviations and the non-equilibrium stress:
    // sum c c f~ = sum c c f - ref cs2 I and sum c c f~eq = dr cs2 I + rho u u.
    R pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int i = 1; i < Q; i++) {
        const R v = f[i];
        if (cx(i)) { pxx += v; jx += (R)cx(i) * v; }
        if (cy(i)) { pyy += v; jy += (R)cy(i) * v; }
        if (cz(i)) { pzz += v; jz += (R)cz(i) * v; }
        if (cx(i) * cy(i)) pxy += (R)(cx(i) * cy(i)) * v;
        if (cx(i) * cz(i)) pxz += (R)(cx(i) * cz(i)) * v;
        if (cy(i) * cz(i)) pyz += (R)(cy(i) * cz(i)) * v;
    }
    const R third = (R)(1.0 / 3.0);
    const R dc = dr * third;
    pxx -= dc + rho * ux * ux;
    pyy -= dc + rho * uy * uy;
    pzz -= dc + rho * uz * uz;
    pxy -= rho * ux * uy;
    pxz -= rho * ux * uz;
    pyz -= rho * uy * uz;
    // momentum non-equilibrium (-F/2 analytically)
    jx -= rho * ux;
    jy -= rho * uy;
    jz -= rho * uz;
    const R tr = pxx + pyy + pzz;
    const R t3 = tr * third;
    const R txx = pxx - t3, tyy = pyy - t3, tzz = pzz - t3;
    // second-order forcing tensor uF + Fu, traceless part and trace
    const R ftr = (R)2 * (ux * Fx + uy * Fy + uz * Fz);
    const R f3 = ftr * third;
    const R gxx = (R)2 * ux * Fx - f3, gyy = (R)2 * uy * Fy - f3, gzz = (R)2 * uz * Fz - f3;
    const R fxy = ux * Fy + uy * Fx, fxz = ux * Fz + uz * Fx, fyz = uy * Fz + uz * Fy;
    const R a_nu = (R)1 - s_nu, a_b = (R)1 - s_b, a_o = (R)1 - s_o, a_e = (R)1 - s_e;
    const R b_nu = (R)1 - (R)0.5 * s_nu, b_b = (R)1 - (R)0.5 * s_b;
    const R uu = ux * ux + uy * uy + uz * uz;
    const R keep = (R)1 - ns;
    // rest population (even only)
    {
        const R w = (R)wgt<Q>(0);
        const R feq = w * (dr - (R)1.5 * rho * uu);
        const R bk = (R)-1.5 * w * tr;
        const R G = f[0] - feq - bk;
        const R post = feq + a_b * bk + a_e * G + b_b * ((R)-1.5 * w * ftr);
        f[0] = keep * post + ns * f[0];
    }
    // opposite pairs: even part E and odd part O, f*_i = E + O, f*_inv(i) = E - O
#pragma unroll
    for (int i = 1; i < Q; i += 2) {
        const int k = i + 1;  // inv(i)
        const R w = (R)wgt<Q>(i);
        const int X = cx(i), Y = cy(i), Z = cz(i);
        const R cu = X * ux + Y * uy + Z * uz;
        const R fe = w * (dr + rho * ((R)4.5 * cu * cu - (R)1.5 * uu));  // even part of f~eq
        const R fo = w * rho * (R)3 * cu;                                 // odd part of f~eq
        const R cTc = (R)(X * X) * txx + (R)(Y * Y) * tyy + (R)(Z * Z) * tzz +
                      (R)(2 * X * Y) * pxy + (R)(2 * X * Z) * pxz + (R)(2 * Y * Z) * pyz;
        const R cGc = (R)(X * X) * gxx + (R)(Y * Y) * gyy + (R)(Z * Z) * gzz +
                      (R)(2 * X * Y) * fxy + (R)(2 * X * Z) * fxz + (R)(2 * Y * Z) * fyz;
        const R sh = (R)4.5 * w * cTc;
        const R bk = (R)1.5 * w * (R)(csq(i) - 1) * tr;
        const R mo = (R)3 * w * (X * jx + Y * jy + Z * jz);
        const R fi = f[i], fk = f[k];
        const R ge = (R)0.5 * (fi + fk) - fe - sh - bk;
        const R go = (R)0.5 * (fi - fk) - fo - mo;
        const R E = fe + a_nu * sh + a_b * bk + a_e * ge + b_nu * ((R)4.5 * w * cGc) +
                    b_b * ((R)1.5 * w * (R)(csq(i) - 1) * ftr);
        const R O = fo + (R)1.5 * w * (X * Fx + Y * Fy + Z * Fz) + a_o * go;
        f[i] = keep * (E + O) + ns * fk;
        f[k] = keep * (E - O) + ns * fi;
    }
}

// Isotropic LB-weighted gradient of the stencil (D3Q19 weights over cs^2 = 1/3), summed over
// differences to the centre value (sum_i c_i w_i = 0): in the bulk, where phi is near +-1, the
// terms are then small instead of cancelling O(1) values, which keeps the FP32 gradient (and
// the direction of the anti-trapping current taken from it) accurate.
template <class R>
__device__ inline void stencil_gradient(const R (&P)[27], R (&gr)[3]) {
    gr[0] = gr[1] = gr[2] = 0;
    const R p0 = P[13];
#pragma unroll
    for (int i = 1; i < 19; i++) {
        const R v = (R)(3.0 * wgt<19>(i)) * (P[(cx(i) + 1) * 9 + (cy(i) + 1) * 3 + (cz(i) + 1)] - p0);
        if (cx(i)) gr[0] += (R)cx(i) * v;
        if (cy(i)) gr[1] += (R)cy(i) * v;
        if (cz(i)) gr[2] += (R)cz(i) * v;
    }
}

// ------------------------------------------------------------------ phase-field source
template <class R>
__device__ inline R pow4(R v) {
    R s = v * v;
    return s * s;
}

// Normal component of J = a_s^2 g + |g|^2 a_s da_s/dg on a face with gradient g (lab frame);
// M (lab <- crystal, row-major) may be null.
// Anisotropy a_s(n) and its derivative d a_s / d n (n the unit normal, crystal frame) for the
// forms MKG_ANISO_CUBIC / CBC / HCP / FACET of mkg.h; e4 = eps4, p0 / p1 = the form's two
// parameters.  Kernel arguments are uniform, so the switch does not diverge.
template <class R>
struct Aniso {
    int kind;
    R e4, p0, p1;
};
template <class R>
__host__ __device__ inline R aniso_eval(const Aniso<R>& an, R n0, R n1, R n2, R& d0, R& d1,
                                        R& d2) {
    const R q0 = n0 * n0, q1 = n1 * n1, q2 = n2 * n2;
    if (an.kind == 1) {  // two cubic harmonics
        const R Q = q0 * q0 + q1 * q1 + q2 * q2, S = q0 * q1 * q2;
        const R c4 = (R)4 * an.p0 + (R)12 * an.p1, c6 = (R)132 * an.p1;
        d0 = c4 * q0 * n0 + c6 * n0 * q1 * q2;
        d1 = c4 * q1 * n1 + c6 * n1 * q0 * q2;
        d2 = c4 * q2 * n2 + c6 * n2 * q0 * q1;
        return (R)1 + an.p0 * (Q - (R)0.6) + an.p1 * ((R)3 * Q + (R)66 * S - (R)(17.0 / 7.0));
    }
    if (an.kind == 2) {  // hexagonal: Re (nx + i ny)^6 and the c axis
        const R H = q0 * q0 * q0 - (R)15 * q0 * q0 * q1 + (R)15 * q0 * q1 * q1 - q1 * q1 * q1;
        d0 = an.p0 * ((R)6 * q0 * q0 * n0 - (R)60 * q0 * n0 * q1 + (R)30 * n0 * q1 * q1);
        d1 = an.p0 * ((R)-30 * q0 * q0 * n1 + (R)60 * q0 * q1 * n1 - (R)6 * q1 * q1 * n1);
        d2 = an.p1 * (R)3 * n2;
        return (R)1 + an.p0 * H + an.p1 * ((R)1.5 * q2 - (R)0.5);
    }
    if (an.kind == 3) {  // facets: |n_k| rounded over nu
        const R v2 = an.p1 * an.p1;
        const R r0 = sqrt(q0 + v2), r1 = sqrt(q1 + v2), r2 = sqrt(q2 + v2);
        d0 = an.p0 * n0 / r0;
        d1 = an.p0 * n1 / r1;
        d2 = an.p0 * n2 / r2;
        return (R)1 + an.p0 * (r0 + r1 + r2 - sqrt((R)1 + v2) - (R)2 * an.p1);
    }
    const R k = (R)16 * an.e4;  // Karma-Rappel cubic
    d0 = k * q0 * n0;
    d1 = k * q1 * n1;
    d2 = k * q2 * n2;
    return (R)1 - (R)3 * an.e4 + (R)4 * an.e4 * (q0 * q0 + q1 * q1 + q2 * q2);
}

// |v| and the unit vector v / |v| from s = |v|^2 > 1e-30 (a normal number).  FP32: one hardware
// reciprocal square root (about 1 ulp), where the fast square root and divisions carry
// range-scaling code; FP64: the exact square root and divisions.
template <class R>
__device__ inline R unit3(R s, R x, R y, R z, R& ux, R& uy, R& uz) {
    if constexpr (sizeof(R) == 4) {
        const R r = __builtin_amdgcn_rsqf(s);
        ux = x * r;
        uy = y * r;
        uz = z * r;
        return s * r;
    } else {
        const R n = sqrt(s);
        ux = x / n;
        uy = y / n;
        uz = z / n;
        return n;
    }
}
// 1 / sqrt(s) for s > 1e-30 (FP32: the hardware reciprocal square root).
template <class R>
__device__ inline R inv_sqrt(R s) {
    if constexpr (sizeof(R) == 4)
        return __builtin_amdgcn_rsqf(s);
    else
        return (R)1 / sqrt(s);
}

// Normal component along `axis` of J = dE/dG, E = |G|^2 a_s(G/|G|)^2 / 2:
//   J = a^2 G + |G| a (D - n (n.D)),  D = d a_s / d n,
// evaluated from a lab-frame face gradient g (crystal frame through M when rot).
template <class R>
__device__ inline R face_flux(R g0, R g1, R g2, int axis, const Aniso<R>& an, const R (&M)[9],
                              bool rot) {
    R c0 = g0, c1 = g1, c2 = g2;
    if (rot) {
        c0 = M[0] * g0 + M[3] * g1 + M[6] * g2;
        c1 = M[1] * g0 + M[4] * g1 + M[7] * g2;
        c2 = M[2] * g0 + M[5] * g1 + M[8] * g2;
    }
    const R g2s = c0 * c0 + c1 * c1 + c2 * c2;
    const R ga = axis == 0 ? g0 : (axis == 1 ? g1 : g2);
    if (!(g2s > (R)1e-30)) return ga;
    R n0, n1, n2;
    const R gn = unit3<R>(g2s, c0, c1, c2, n0, n1, n2);
    R D0, D1, D2;
    const R as = aniso_eval<R>(an, n0, n1, n2, D0, D1, D2);
    const R nD = n0 * D0 + n1 * D1 + n2 * D2;
    const R k = as * gn;
    const R N0 = k * (D0 - n0 * nD);
    const R N1 = k * (D1 - n1 * nD);
    const R N2 = k * (D2 - n2 * nD);
    R Nl;
    if (rot)
        Nl = M[axis * 3 + 0] * N0 + M[axis * 3 + 1] * N1 + M[axis * 3 + 2] * N2;
    else
        Nl = axis == 0 ? N0 : (axis == 1 ? N1 : N2);
    return as * as * ga + Nl;
}

template <class R>
__device__ inline R phase_source(const R (&P)[27], R upf, const Aniso<R>& an, R lambda, R pf_dt,
                                 R pf_wdx2, const R (&M)[9], bool rot, R& a_out) {
#define PP(dx, dy, dz) P[((dx) + 1) * 9 + ((dy) + 1) * 3 + ((dz) + 1)]
    R divJ = 0, lap = 0;
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
#pragma unroll
        for (int sd = 0; sd < 2; sd++) {
            const int side = sd == 0 ? 1 : -1;
            const int o0 = ax == 0 ? side : 0, o1 = ax == 1 ? side : 0, o2 = ax == 2 ? side : 0;
            R g[3];
#pragma unroll
            for (int b = 0; b < 3; b++) {
                if (b == ax) {
                    g[b] = (R)side * (PP(o0, o1, o2) - PP(0, 0, 0));
                } else {
                    const int e0 = b == 0, e1 = b == 1, e2 = b == 2;
                    g[b] = (R)0.25 * (PP(e0, e1, e2) - PP(-e0, -e1, -e2) +
                                      PP(o0 + e0, o1 + e1, o2 + e2) -
                                      PP(o0 - e0, o1 - e1, o2 - e2));
                }
            }
            const R J = face_flux<R>(g[0], g[1], g[2], ax, an, M, rot);
            divJ += (R)side * J;
            lap += (R)side * g[ax];
        }
    }
    // isotropic LB-weighted cell gradient (D3Q19 weights over cs^2 = 1/3), summed over
    // differences to the centre value (no cancellation of O(1) terms in the bulk)
    R gr0 = 0, gr1 = 0, gr2 = 0;
#pragma unroll
    for (int i = 1; i < 19; i++) {
        const R v = (R)(3.0 * wgt<19>(i)) * (PP(cx(i), cy(i), cz(i)) - PP(0, 0, 0));
        if (cx(i)) gr0 += (R)cx(i) * v;
        if (cy(i)) gr1 += (R)cy(i) * v;
        if (cz(i)) gr2 += (R)cz(i) * v;
    }
    R c0 = gr0, c1 = gr1, c2 = gr2;
    if (rot) {
        c0 = M[0] * gr0 + M[3] * gr1 + M[6] * gr2;
        c1 = M[1] * gr0 + M[4] * gr1 + M[7] * gr2;
        c2 = M[2] * gr0 + M[5] * gr1 + M[8] * gr2;
    }
    R as = 1;
    const R gg = c0 * c0 + c1 * c1 + c2 * c2;
    if (gg > (R)1e-30) {
        R n0, n1, n2, D0, D1, D2;
        unit3<R>(gg, c0, c1, c2, n0, n1, n2);
        as = aniso_eval<R>(an, n0, n1, n2, D0, D1, D2);
    }
    const R ph = PP(0, 0, 0);
    const R d = (R)1 - ph * ph;
    const R Qs = (ph - lambda * upf * d) * d;
    const R a2 = as * as;
    a_out = as;
    return pf_dt / a2 * ((divJ - a2 * lap) * pf_wdx2 + Qs);
#undef PP
}

// reset_flags (MaterKit/inc/mkPhaseField.h:37); the Allen-Cahn model keeps every flag.
__device__ inline int8_t classify(int8_t old, double phi, double phi_s, int model = 0) {
    if (old == kIsWalls || old < 0 || model != 0) return old;
    if (phi >= phi_s) return (int8_t)kIsSolid;
    if (phi > -phi_s) return (int8_t)kIsInter;
    return (int8_t)kIsFluid;
}

}  // namespace mkg
