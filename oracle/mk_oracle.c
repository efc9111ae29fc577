/*
 * mk_oracle.c -- FP64 CPU oracle (TEST INFRASTRUCTURE ONLY; see mk_oracle.h).
 *
 * Written as a direct, unoptimised restatement: every step
 *   1. (phase advanced)  phi_prev <- phi, phi <- sum_i g_i, flags re-classified from phi
 *      (mkFluidsUser::relaxation first loop, MaterKit/src/mkFluidsUser.cpp:147-172, and
 *       mkPhaseField::reset_flags, MaterKit/inc/mkPhaseField.h:37)
 *   2. macroscopic moments of every present lattice at time level n
 *   3. collision of every advanced lattice (post-collision buffers)
 *   4. pull streaming with face rules and half-way bounce-back at wall cells
 *      (pull form and "Plan A" bounce-back of mkFluidsUser::propagate,
 *       MaterKit/src/mkFluidsUser.cpp:68-134)
 *
 * Published methods used (the reference's own bodies are in a closed DLL):
 *   - second-order equilibrium, Guo-Zheng-Shi (2002) forcing with u = (sum c f + F/2)/rho
 *   - TRT (Ginzburg 2008) with the magic parameter Lambda
 *   - MRT as relaxation of w-orthogonal moment subspaces (mass, momentum, traceless shear,
 *     bulk, odd ghosts, even ghosts), projectors built numerically from the velocity set
 *   - partial ("porous") bounce-back of Walsh, Burwinkle & Saar (2009) for the solid fraction
 *   - Karma-Rappel (1998) anisotropic phase field, cubic a_s(n) = 1 - 3 eps4 + 4 eps4 sum n^4,
 *     solved as a D3Q19/D3Q7 LB diffusion (omega_g = 1/(dt/dx^2/cs^2 + 1/2), the form at
 *     MaterKit/src/mkFluidsUser.cpp:282) plus the source w_i * S, where S carries the
 *     anisotropic face-flux divergence and Q = (phi - lambda u (1-phi^2))(1-phi^2)
 *     (MaterKit/src/mkFluidsUser.cpp:261-264); the melt carries phi with the flux
 *     (phi + 1) u, so the bulk liquid phi = -1 stays exact where div u != 0
 *   - D3Q7 advection-diffusion for T and for the dilute-alloy variable c = A(phi) U
 *     (Karma 2001, Echebarria et al. 2004), latent/rejection source w_i L dphi; the solute
 *     flux D q grad U is split into diffusion of c with D q / A and a drift velocity
 *     V = D q grad A / A^2 carried by the equilibrium
 */
#include "mk_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Planted mistakes.  The default build has ORC_MUTANT 0.  The CPU tests build copies of this
 * file with ORC_MUTANT set to one of the values below (oracle.planted) and check that the pin
 * written for that term fails on the copy:
 *   1  D3Q7 thermal advective equilibrium with cs^2 = 1/3 instead of 1/4
 *   2  latent-heat source with its sign flipped
 *   3  latent-heat source with twice its factor
 *   4  Boussinesq force with twice its factor
 *   5  Boussinesq force with its sign flipped
 *   6  D3Q7 solute advective equilibrium with cs^2 = 1/3
 *   7  solute rejection source with its sign flipped
 *   8  velocity face with f_eq(rho_n, u_n) - f_eq(rho_n, u_b) (the difference reversed) */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

/* ------------------------------------------------------------------ lattices */
/* Direction order: rest, 6 faces (+x,-x,+y,-y,+z,-z), 12 edges, 8 corners; opposite pairs
 * are (2k-1, 2k).  x is the slowest axis of the linear index (x*ny + y)*nz + z. */
static const int E27[27][3] = {
    {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},  {0, 0, -1},
    {1, 1, 0},   {-1, -1, 0}, {1, -1, 0},  {-1, 1, 0},  {1, 0, 1},   {-1, 0, -1}, {1, 0, -1},
    {-1, 0, 1},  {0, 1, 1},   {0, -1, -1}, {0, 1, -1},  {0, -1, 1},  {1, 1, 1},  {-1, -1, -1},
    {1, 1, -1},  {-1, -1, 1}, {1, -1, 1},  {-1, 1, -1}, {-1, 1, 1},  {1, -1, -1}};

typedef struct {
    int q;
    double w[27];
    double cs2;
} lat_t;

static lat_t make_lat(int q) {
    lat_t L;
    memset(&L, 0, sizeof(L));
    L.q = q;
    for (int i = 0; i < q; i++) {
        int n = abs(E27[i][0]) + abs(E27[i][1]) + abs(E27[i][2]);
        if (q == 7) L.w[i] = (n == 0) ? 0.25 : 0.125;
        if (q == 19) L.w[i] = (n == 0) ? 1.0 / 3.0 : (n == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
        if (q == 27)
            L.w[i] = (n == 0) ? 8.0 / 27.0
                              : (n == 1 ? 2.0 / 27.0 : (n == 2 ? 1.0 / 54.0 : 1.0 / 216.0));
    }
    L.cs2 = (q == 7) ? 0.25 : 1.0 / 3.0;
    return L;
}

static int inv_dir(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }

int orc_lattice(int q, int* e, double* w, int* inv, double* cs2) {
    if (q != 7 && q != 19 && q != 27) return -1;
    lat_t L = make_lat(q);
    for (int i = 0; i < q; i++) {
        for (int d = 0; d < 3; d++) e[3 * i + d] = E27[i][d];
        w[i] = L.w[i];
        inv[i] = inv_dir(i);
    }
    *cs2 = L.cs2;
    return 0;
}

size_t orc_params_size(void) { return sizeof(orc_params); }

/* ------------------------------------------------------------------ RNG */
uint64_t orc_splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double orc_uniform(uint64_t* s) { return (double)(orc_splitmix64(s) >> 11) * 0x1.0p-53; }
double orc_cell_uniform(uint64_t seed, uint64_t cell) {
    uint64_t s = seed ^ (cell * 0xD1B54A32D192ED03ULL);
    return orc_uniform(&s);
}

/* ------------------------------------------------------------------ helpers */
#define NCELL(p) ((long)(p)->nx * (p)->ny * (p)->nz)
static long lin(const orc_params* p, int x, int y, int z) {
    return ((long)x * p->ny + y) * p->nz + z;
}

/* Resolve a coordinate against the domain.  Returns the face (0..5) through which the point
 * leaves a non-periodic side, or -1 (point inside, periodic sides wrapped in place).  Axes
 * are examined in the order x, y, z. */
static int resolve(const orc_params* p, const int* bc, int* X, int* Y, int* Z) {
    int dims[3] = {p->nx, p->ny, p->nz};
    int* c[3] = {X, Y, Z};
    for (int a = 0; a < 3; a++) {
        int v = *c[a];
        if (v < 0 || v >= dims[a]) {
            int face = 2 * a + (v < 0 ? 0 : 1);
            if (bc[face] == 0) {
                *c[a] = (v + dims[a]) % dims[a];
            } else {
                return face;
            }
        }
    }
    return -1;
}

/* Phase-field stencil neighbour: periodic sides wrap, other sides replicate the edge. */
static int clampwrap(int v, int n, int periodic) {
    if (v >= 0 && v < n) return v;
    if (periodic) return (v + n) % n;
    return v < 0 ? 0 : n - 1;
}

/* ------------------------------------------------------------------ equilibria */
void orc_equilibrium(int q, double rho, const double* u, double* feq) {
    lat_t L = make_lat(q);
    double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    for (int i = 0; i < q; i++) {
        double cu = E27[i][0] * u[0] + E27[i][1] * u[1] + E27[i][2] * u[2];
        feq[i] = L.w[i] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
    }
}

/* He & Luo (1997) incompressible equilibrium: w_i [rho + rho0 (3 c.u + 4.5 (c.u)^2 - 1.5 u.u)]
 * (the getEquPDFsHe variant, MaterKit/inc/mkNode.h:34-50) */
void orc_equilibrium_he(int q, double rho, double rho0, const double* u, double* feq) {
    lat_t L = make_lat(q);
    double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    for (int i = 0; i < q; i++) {
        double cu = E27[i][0] * u[0] + E27[i][1] * u[1] + E27[i][2] * u[2];
        feq[i] = L.w[i] * (rho + rho0 * (3.0 * cu + 4.5 * cu * cu - 1.5 * uu));
    }
}

/* velocity of a cell: (j + F/2) / rho, or / rho0 for the incompressible equilibrium */
static double flow_den(const orc_params* p, double rho) { return p->flow_incomp ? p->rho0 : rho; }

/* Guo forcing term without its rate factor: w_i [3 (c-u).F + 9 (c.u)(c.F)] */
static void guo_term(int q, const double* u, const double* F, double* Fi) {
    lat_t L = make_lat(q);
    for (int i = 0; i < q; i++) {
        double cu = E27[i][0] * u[0] + E27[i][1] * u[1] + E27[i][2] * u[2];
        double cF = E27[i][0] * F[0] + E27[i][1] * F[1] + E27[i][2] * F[2];
        double uF = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
        Fi[i] = L.w[i] * (3.0 * (cF - uF) + 9.0 * cu * cF);
    }
}

/* ------------------------------------------------------------------ MRT projectors */
/* Projector onto span(B) in the inner product <a,b> = sum a_i b_i / w_i:
 * P = B (B^T W^-1 B)^-1 B^T W^-1, B is q x m. */
static void projector(int q, const double* w, const double* B, int m, double* P) {
    double G[27 * 27], Gi[27 * 27];
    for (int a = 0; a < m; a++)
        for (int b = 0; b < m; b++) {
            double s = 0;
            for (int i = 0; i < q; i++) s += B[i * m + a] * B[i * m + b] / w[i];
            G[a * m + b] = s;
        }
    /* Gauss-Jordan inverse of the m x m Gram matrix */
    for (int a = 0; a < m; a++)
        for (int b = 0; b < m; b++) Gi[a * m + b] = (a == b) ? 1.0 : 0.0;
    for (int col = 0; col < m; col++) {
        int piv = col;
        for (int r = col + 1; r < m; r++)
            if (fabs(G[r * m + col]) > fabs(G[piv * m + col])) piv = r;
        for (int k = 0; k < m; k++) {
            double t = G[col * m + k];
            G[col * m + k] = G[piv * m + k];
            G[piv * m + k] = t;
            t = Gi[col * m + k];
            Gi[col * m + k] = Gi[piv * m + k];
            Gi[piv * m + k] = t;
        }
        double d = G[col * m + col];
        for (int k = 0; k < m; k++) {
            G[col * m + k] /= d;
            Gi[col * m + k] /= d;
        }
        for (int r = 0; r < m; r++)
            if (r != col) {
                double f = G[r * m + col];
                for (int k = 0; k < m; k++) {
                    G[r * m + k] -= f * G[col * m + k];
                    Gi[r * m + k] -= f * Gi[col * m + k];
                }
            }
    }
    for (int i = 0; i < q; i++)
        for (int j = 0; j < q; j++) {
            double s = 0;
            for (int a = 0; a < m; a++)
                for (int b = 0; b < m; b++) s += B[i * m + a] * Gi[a * m + b] * B[j * m + b];
            P[i * q + j] = s / w[j];
        }
}

typedef struct {
    int built;
    double mom[27 * 27], shear[27 * 27], bulk[27 * 27], godd[27 * 27], geven[27 * 27];
} mrt_t;
static mrt_t MRT19, MRT27;

static void build_mrt(int q, mrt_t* M) {
    if (M->built) return;
    lat_t L = make_lat(q);
    double B[27 * 5] = {0}, Pm[27 * 27];
    /* mass */
    for (int i = 0; i < q; i++) B[i] = L.w[i];
    projector(q, L.w, B, 1, Pm);
    /* momentum */
    for (int i = 0; i < q; i++)
        for (int a = 0; a < 3; a++) B[i * 3 + a] = L.w[i] * E27[i][a];
    projector(q, L.w, B, 3, M->mom);
    /* traceless shear: xy, xz, yz, xx-yy, yy-zz */
    for (int i = 0; i < q; i++) {
        const int* c = E27[i];
        B[i * 5 + 0] = L.w[i] * c[0] * c[1];
        B[i * 5 + 1] = L.w[i] * c[0] * c[2];
        B[i * 5 + 2] = L.w[i] * c[1] * c[2];
        B[i * 5 + 3] = L.w[i] * (c[0] * c[0] - c[1] * c[1]);
        B[i * 5 + 4] = L.w[i] * (c[1] * c[1] - c[2] * c[2]);
    }
    projector(q, L.w, B, 5, M->shear);
    /* bulk */
    for (int i = 0; i < q; i++) {
        const int* c = E27[i];
        B[i] = L.w[i] * (c[0] * c[0] + c[1] * c[1] + c[2] * c[2] - 3.0 * L.cs2);
    }
    projector(q, L.w, B, 1, M->bulk);
    /* ghosts = complement, split by parity under inversion */
    for (int i = 0; i < q; i++)
        for (int j = 0; j < q; j++) {
            double id = (i == j) ? 1.0 : 0.0;
            double g = id - Pm[i * q + j] - M->mom[i * q + j] - M->shear[i * q + j] -
                       M->bulk[i * q + j];
            /* P_ghost * (I -/+ R)/2 : column j of R is e_{inv(j)} */
            M->godd[i * q + j] = 0.5 * g;
            M->geven[i * q + j] = 0.5 * g;
        }
    double tmp_odd[27 * 27], tmp_even[27 * 27];
    for (int i = 0; i < q; i++)
        for (int j = 0; j < q; j++) {
            int jr = inv_dir(j);
            tmp_odd[i * q + j] = M->godd[i * q + j] - M->godd[i * q + jr];
            tmp_even[i * q + j] = M->geven[i * q + j] + M->geven[i * q + jr];
        }
    memcpy(M->godd, tmp_odd, sizeof(tmp_odd));
    memcpy(M->geven, tmp_even, sizeof(tmp_even));
    M->built = 1;
}

static void matvec(int q, const double* P, const double* v, double* out) {
    for (int i = 0; i < q; i++) {
        double s = 0;
        for (int j = 0; j < q; j++) s += P[i * q + j] * v[j];
        out[i] = s;
    }
}

/* ------------------------------------------------------------------ flow collision */
/* f: pre-collision populations of one cell, F: force density.  out: post-collision.
 * ns: partial bounce-back solid fraction (Walsh et al. 2009). */
void orc_collide_cell(const orc_params* p, int op, const double* f, const double* F, double* out,
                      double ns) {
    int q = p->flow_q;
    double rho = 0, j[3] = {0, 0, 0};
    for (int i = 0; i < q; i++) {
        rho += f[i];
        for (int a = 0; a < 3; a++) j[a] += E27[i][a] * f[i];
    }
    double u[3];
    for (int a = 0; a < 3; a++) u[a] = (j[a] + 0.5 * F[a]) / flow_den(p, rho);
    double feq[27], Fi[27], fs[27];
    if (p->flow_incomp)
        orc_equilibrium_he(q, rho, p->rho0, u, feq);
    else
        orc_equilibrium(q, rho, u, feq);
    guo_term(q, u, F, Fi);
    if (op == ORC_OP_BGK) {
        double om = p->s_nu;
        for (int i = 0; i < q; i++) fs[i] = f[i] - om * (f[i] - feq[i]) + (1.0 - 0.5 * om) * Fi[i];
    } else if (op == ORC_OP_TRT) {
        double wp = p->s_nu, wm = p->s_odd;
        for (int i = 0; i < q; i++) {
            int k = inv_dir(i);
            double fp = 0.5 * (f[i] + f[k]), fm = 0.5 * (f[i] - f[k]);
            double ep = 0.5 * (feq[i] + feq[k]), em = 0.5 * (feq[i] - feq[k]);
            double Fp = 0.5 * (Fi[i] + Fi[k]), Fm = 0.5 * (Fi[i] - Fi[k]);
            fs[i] = f[i] - wp * (fp - ep) - wm * (fm - em) + (1.0 - 0.5 * wp) * Fp +
                    (1.0 - 0.5 * wm) * Fm;
        }
    } else {
        mrt_t* M = (q == 19) ? &MRT19 : &MRT27;
        build_mrt(q, M);
        double neq[27], t[27];
        for (int i = 0; i < q; i++) {
            neq[i] = f[i] - feq[i];
            fs[i] = feq[i];
        }
        const double* Ps[5] = {M->mom, M->shear, M->bulk, M->godd, M->geven};
        double rates[5] = {1.0, p->s_nu, p->s_bulk, p->s_odd, p->s_even};
        for (int k = 0; k < 5; k++) {
            matvec(q, Ps[k], neq, t);
            for (int i = 0; i < q; i++) fs[i] += (1.0 - rates[k]) * t[i];
            matvec(q, Ps[k], Fi, t);
            for (int i = 0; i < q; i++) fs[i] += (1.0 - 0.5 * rates[k]) * t[i];
        }
    }
    for (int i = 0; i < q; i++) out[i] = (1.0 - ns) * fs[i] + ns * f[inv_dir(i)];
}

/* ------------------------------------------------------------------ phase field */
/* a_s(n) and D = d a_s / d n for the unit crystal-frame normal n, by form (aniso_kind):
 *   0 cubic (Karma-Rappel 1998): 1 - 3 e4 + 4 e4 sum n^4
 *   1 two cubic harmonics (Haxhimali et al. 2006): 1 + e1 (Q - 3/5) + e2 (3Q + 66S - 17/7),
 *     Q = sum n^4, S = nx^2 ny^2 nz^2
 *   2 hexagonal: 1 + e6 Re (nx + i ny)^6 + ez (3 nz^2 - 1)/2
 *   3 faceted: 1 + ef (sum sqrt(n_k^2 + nu^2) - sqrt(1 + nu^2) - 2 nu)
 * (the cubic, cbc / hcp and facet variants of mkPhaseField, MaterKit/inc/mkPhaseField.h:49-64,
 * whose bodies are not public: these are the published forms named above) */
double orc_aniso(const orc_params* p, const double* n, double* D) {
    double x = n[0], y = n[1], z = n[2];
    double x2 = x * x, y2 = y * y, z2 = z * z;
    double e1 = p->aniso[0], e2 = p->aniso[1];
    switch (p->aniso_kind) {
        case 1: {
            double Q = x2 * x2 + y2 * y2 + z2 * z2, S = x2 * y2 * z2;
            D[0] = 4.0 * e1 * x2 * x + e2 * (12.0 * x2 * x + 132.0 * x * y2 * z2);
            D[1] = 4.0 * e1 * y2 * y + e2 * (12.0 * y2 * y + 132.0 * y * x2 * z2);
            D[2] = 4.0 * e1 * z2 * z + e2 * (12.0 * z2 * z + 132.0 * z * x2 * y2);
            return 1.0 + e1 * (Q - 0.6) + e2 * (3.0 * Q + 66.0 * S - 17.0 / 7.0);
        }
        case 2: {
            double H = x2 * x2 * x2 - 15.0 * x2 * x2 * y2 + 15.0 * x2 * y2 * y2 - y2 * y2 * y2;
            D[0] = e1 * (6.0 * x2 * x2 * x - 60.0 * x2 * x * y2 + 30.0 * x * y2 * y2);
            D[1] = e1 * (-30.0 * x2 * x2 * y + 60.0 * x2 * y2 * y - 6.0 * y2 * y2 * y);
            D[2] = e2 * 3.0 * z;
            return 1.0 + e1 * H + e2 * (1.5 * z2 - 0.5);
        }
        case 3: {
            double v2 = e2 * e2;
            double r[3] = {sqrt(x2 + v2), sqrt(y2 + v2), sqrt(z2 + v2)};
            for (int k = 0; k < 3; k++) D[k] = e1 * n[k] / r[k];
            return 1.0 + e1 * (r[0] + r[1] + r[2] - sqrt(1.0 + v2) - 2.0 * e2);
        }
        default: {
            double e4 = p->eps4;
            for (int k = 0; k < 3; k++) D[k] = 16.0 * e4 * n[k] * n[k] * n[k];
            return 1.0 - 3.0 * e4 + 4.0 * e4 * (x2 * x2 + y2 * y2 + z2 * z2);
        }
    }
}

/* Normal component along `axis` of J = a_s^2 grad phi + |grad phi|^2 a_s d a_s / d grad phi
 * = a^2 G + |G| a (D - n (n.D)), evaluated from a face gradient g (lab frame); R (lab <- crystal)
 * may be NULL. */
static double face_flux(const orc_params* p, const double* R, const double* g, int axis) {
    double gc[3];
    if (R) {
        for (int a = 0; a < 3; a++) gc[a] = R[0 * 3 + a] * g[0] + R[1 * 3 + a] * g[1] + R[2 * 3 + a] * g[2];
    } else {
        gc[0] = g[0];
        gc[1] = g[1];
        gc[2] = g[2];
    }
    double g2 = gc[0] * gc[0] + gc[1] * gc[1] + gc[2] * gc[2];
    if (g2 <= 1e-30) return g[axis];
    double gn = sqrt(g2);
    double n[3] = {gc[0] / gn, gc[1] / gn, gc[2] / gn};
    double D[3];
    double as = orc_aniso(p, n, D);
    double nD = n[0] * D[0] + n[1] * D[1] + n[2] * D[2];
    double Nc[3];
    for (int a = 0; a < 3; a++) Nc[a] = as * gn * (D[a] - n[a] * nD);
    double Nl = R ? (R[axis * 3 + 0] * Nc[0] + R[axis * 3 + 1] * Nc[1] + R[axis * 3 + 2] * Nc[2])
                  : Nc[axis];
    return as * as * g[axis] + Nl;
}

/* Exposed for tests: normal flux component of a face with gradient g, grain-less when R is
 * NULL. */
double orc_face_flux(const orc_params* p, const double* R, const double* g, int axis) {
    return face_flux(p, R, g, axis);
}

#define P27(dx, dy, dz) phi27[((dx) + 1) * 9 + ((dy) + 1) * 3 + ((dz) + 1)]

/* Source S of the phase-field LB equation for one cell from its 3x3x3 phi neighbourhood.
 * S = dt/a^2 [ (div J - a^2 lap phi) W0^2/dx^2 + Q ],  Q = (phi - lambda u (1-phi^2))(1-phi^2)
 * with J on the six faces (normal difference across the face, transverse central differences
 * averaged over the two cells), lap the matching 7-point Laplacian, and a = a_s of the
 * isotropic LB-weighted cell gradient (compute_normal_LB form, MaterKit/src/mkFluidsUser.cpp:
 * 211-217). */
void orc_phase_source(const orc_params* p, const double* phi27, double upf, int grain,
                      double* S_out, double* a_out) {
    const double* R = (p->ngrains > 0 && grain > 0) ? p->rot[grain - 1] : NULL;
    double divJ = 0, lap = 0;
    for (int ax = 0; ax < 3; ax++) {
        for (int side = 1; side >= -1; side -= 2) {
            int o[3] = {0, 0, 0};
            o[ax] = side;
            double g[3];
            for (int b = 0; b < 3; b++) {
                if (b == ax) {
                    g[b] = side * (P27(o[0], o[1], o[2]) - P27(0, 0, 0));
                } else {
                    int e[3] = {0, 0, 0};
                    e[b] = 1;
                    g[b] = 0.25 * (P27(e[0], e[1], e[2]) - P27(-e[0], -e[1], -e[2]) +
                                   P27(o[0] + e[0], o[1] + e[1], o[2] + e[2]) -
                                   P27(o[0] - e[0], o[1] - e[1], o[2] - e[2]));
                }
            }
            double J = face_flux(p, R, g, ax);
            divJ += side * J;
            lap += side * g[ax];
        }
    }
    /* isotropic LB-weighted cell gradient, D3Q19 weights, / cs^2 */
    lat_t L = make_lat(19);
    double gr[3] = {0, 0, 0};
    for (int i = 1; i < 19; i++)
        for (int a = 0; a < 3; a++) gr[a] += 3.0 * L.w[i] * E27[i][a] * P27(E27[i][0], E27[i][1], E27[i][2]);
    double a = 1.0;
    double grc[3];
    if (R) {
        for (int k = 0; k < 3; k++) grc[k] = R[0 * 3 + k] * gr[0] + R[1 * 3 + k] * gr[1] + R[2 * 3 + k] * gr[2];
    } else {
        grc[0] = gr[0];
        grc[1] = gr[1];
        grc[2] = gr[2];
    }
    double gg = grc[0] * grc[0] + grc[1] * grc[1] + grc[2] * grc[2];
    if (gg > 1e-30) {
        double gn = sqrt(gg);
        double n[3] = {grc[0] / gn, grc[1] / gn, grc[2] / gn};
        double D[3];
        a = orc_aniso(p, n, D);
    }
    double ph = P27(0, 0, 0);
    double d = 1.0 - ph * ph;
    double Q = (ph - p->lambda * upf * d) * d;
    double a2 = a * a;
    *S_out = p->pf_dt / a2 * ((divJ - a2 * lap) * p->pf_wdx2 + Q);
    if (a_out) *a_out = a;
}

/* ------------------------------------------------------------------ moments */
static double solid_fraction(double phi) {
    double fs = 0.5 * (1.0 + phi);
    return fs < 0 ? 0 : (fs > 1 ? 1 : fs);
}

typedef struct {
    double *rho, *u, *T, *C, *U, *ns;
} macro_t;

/* Body force density on the flow: the constant force plus the Boussinesq buoyancy
 * rho0 * buoy * (T - T_ref), buoy = -beta_T g (mk_Gravity, mk_beta_T,
 * MaterKit/inc/mkFluidsLBM.h:123-131). */
static void body_force(const orc_params* p, double T, double* F) {
    double b = p->rho0 * (T - p->T_ref);
#if ORC_MUTANT == 4
    b *= 2.0;
#elif ORC_MUTANT == 5
    b = -b;
#endif
    for (int a = 0; a < 3; a++) {
        F[a] = p->force[a];
        if (p->thermal) F[a] += p->buoy[a] * b;
    }
}

static void compute_macro(const orc_params* p, const orc_state* s, macro_t* m) {
    long N = NCELL(p);
    lat_t Lf = make_lat(p->flow_q ? p->flow_q : 19);
#pragma omp parallel for schedule(static)
    for (long x = 0; x < N; x++) {
        double phi = p->phase_q ? s->phi[x] : -1.0;
        if (p->thermal) {
            double T = 0;
            for (int i = 0; i < 7; i++) T += s->h[i * N + x];
            m->T[x] = T;
        }
        if (p->solute) {
            double C = 0;
            for (int i = 0; i < 7; i++) C += s->c[i * N + x];
            double A = 0.5 * ((1.0 + p->k_part) - (1.0 - p->k_part) * phi);
            m->C[x] = C;
            m->U[x] = C / A;
        }
        double ns = 0;
        if (p->phase_q) {
            if (s->flag[x] == 2) ns = 1.0;
            else if (s->flag[x] == 1) ns = solid_fraction(phi);
        }
        m->ns[x] = ns;
        if (p->flow_q) {
            double rho = 0, j[3] = {0, 0, 0};
            for (int i = 0; i < p->flow_q; i++) {
                double fi = s->f[i * N + x];
                rho += fi;
                for (int a = 0; a < 3; a++) j[a] += E27[i][a] * fi;
            }
            double F[3];
            body_force(p, m->T[x], F);
            m->rho[x] = rho;
            for (int a = 0; a < 3; a++) m->u[3 * x + a] = (j[a] + 0.5 * F[a]) / flow_den(p, rho);
        } else {
            for (int a = 0; a < 3; a++) m->u[3 * x + a] = 0;
        }
        (void)Lf;
    }
}

void orc_moments(const orc_params* p, const orc_state* s, double* rho, double* vel3, double* T,
                 double* C, double* phi_now) {
    long N = NCELL(p);
    macro_t m;
    m.rho = malloc(N * sizeof(double));
    m.u = malloc(3 * N * sizeof(double));
    m.T = malloc(N * sizeof(double));
    m.C = malloc(N * sizeof(double));
    m.U = malloc(N * sizeof(double));
    m.ns = malloc(N * sizeof(double));
    compute_macro(p, s, &m);
    if (rho && p->flow_q) memcpy(rho, m.rho, N * sizeof(double));
    if (vel3) memcpy(vel3, m.u, 3 * N * sizeof(double));
    if (T && p->thermal) memcpy(T, m.T, N * sizeof(double));
    if (C && p->solute) memcpy(C, m.C, N * sizeof(double));
    if (phi_now && p->phase_q) {
        lat_t L = make_lat(p->phase_q);
        for (long x = 0; x < N; x++) {
            double ph = 0;
            for (int i = 0; i < L.q; i++) ph += s->g[i * N + x];
            phi_now[x] = ph;
        }
    }
    free(m.rho); free(m.u); free(m.T); free(m.C); free(m.U); free(m.ns);
}

/* ------------------------------------------------------------------ streaming */
/* post: post-collision populations of the lattice; out: new pre-collision populations.
 * kind 0 flow (bounce with moving-wall term), 1 scalar (noflux / Dirichlet), value per face.
 * wscale (flow, may be NULL = 1): per-cell factor of the moving-wall term, the liquid fraction
 * 1 - n_s, so that a solid cell on an inlet receives no mass.  dscale (scalar, may be NULL =
 * 1): per-cell factor of the Dirichlet value (A(phi) for the solute, whose faces fix U).
 * vel (flow): the cells' fluid velocities, for pressure faces (code 4). */
/* Velocity of the wall of face f next to cell (x, y, z): u_face[f] plus the tangential part
 * of wall_omega x (r - wall_center), r the wall point half a cell through the face
 * (rotating walls, bdyRotation / bdySpiral, MaterKit/inc/mkFluidsLBM.h:64-65). */
static void wall_velocity(const orc_params* p, int f, int x, int y, int z, double* uw) {
    double nrm[3] = {0, 0, 0};
    nrm[f / 2] = (f & 1) ? 1.0 : -1.0;
    double r[3] = {x + 0.5 * nrm[0] - p->wall_center[0], y + 0.5 * nrm[1] - p->wall_center[1],
                   z + 0.5 * nrm[2] - p->wall_center[2]};
    const double* w = p->wall_omega;
    double rot[3] = {w[1] * r[2] - w[2] * r[1], w[2] * r[0] - w[0] * r[2],
                     w[0] * r[1] - w[1] * r[0]};
    double rn = rot[0] * nrm[0] + rot[1] * nrm[1] + rot[2] * nrm[2];
    for (int a = 0; a < 3; a++) uw[a] = p->u_face[f][a] + rot[a] - rn * nrm[a];
}

static void stream(const orc_params* p, const orc_state* s, int q, const int* bc, int kind,
                   const double* fvals, const double* wscale, const double* dscale,
                   const double* vel, const double* post, const double* old, double* out) {
    long N = NCELL(p);
    lat_t L = make_lat(q);
#pragma omp parallel for collapse(2) schedule(static)
    for (int x = 0; x < p->nx; x++)
        for (int y = 0; y < p->ny; y++)
            for (int z = 0; z < p->nz; z++) {
                long id = lin(p, x, y, z);
                if (s->flag[id] == 3) {
                    for (int i = 0; i < q; i++) out[i * N + id] = old[i * N + id];
                    continue;
                }
                for (int i = 0; i < q; i++) {
                    int X = x - E27[i][0], Y = y - E27[i][1], Z = z - E27[i][2];
                    int face = resolve(p, bc, &X, &Y, &Z);
                    int k = inv_dir(i);
                    double v;
                    if (face >= 0) {
                        int t = bc[face];
                        if (t == 3 || (t == 5 && kind == 0)) {
                            /* outflow; the layer of a velocity face continues itself */
                            v = post[i * N + id];
                        } else if (t == 4 && kind == 0) {
                            /* pressure outlet: anti-bounce-back to the equilibrium at rho0 and
                             * the cell's velocity */
                            const double* u = vel + 3 * id;
                            double cu = E27[i][0] * u[0] + E27[i][1] * u[1] + E27[i][2] * u[2];
                            double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
                            v = -post[k * N + id] +
                                2.0 * L.w[i] * p->rho0 * (1.0 + 4.5 * cu * cu - 1.5 * uu);
                        } else if (t == 2 && kind == 1) {
                            double d = dscale ? dscale[id] : 1.0;
                            v = -post[k * N + id] + 2.0 * L.w[i] * d * fvals[face];
                        } else {
                            v = post[k * N + id];
                            if (kind == 0) {
                                double uw[3];
                                wall_velocity(p, face, x, y, z, uw);
                                double cu = E27[i][0] * uw[0] + E27[i][1] * uw[1] + E27[i][2] * uw[2];
                                v += 6.0 * L.w[i] * p->rho0 * cu * (wscale ? wscale[id] : 1.0);
                            }
                        }
                    } else {
                        long src = lin(p, X, Y, Z);
                        if (s->flag[src] == 3) v = post[k * N + id];
                        else v = post[i * N + src];
                    }
                    out[i * N + id] = v;
                }
            }
}

/* ------------------------------------------------------------------ flags */
static int8_t classify(const orc_params* p, int8_t old, double phi) {
    if (old == 3 || old < 0 || p->pf_model != 0) return old; /* Allen-Cahn keeps every flag */
    if (phi >= p->phi_s) return 2;
    if (phi > -p->phi_s) return 1;
    return 0;
}

/* Velocity faces by non-equilibrium extrapolation (mkFluidsUser::boundary_vel,
 * MaterKit/src/mkFluidsUser.cpp:455-505): every non-wall cell b of the face layer takes the
 * post-collision populations of its inward neighbour n (unless n is a wall cell) plus
 * f_eq(rho_n, u_b) - f_eq(rho_n, u_n), with rho_n = sum f*_n, u_n = sum c f*_n / rho_n (:485-500)
 * and u_b the face velocity.  Velocity faces share no edge, so no n is itself replaced. */
static void nee_faces(const orc_params* p, orc_state* s, int q, double* post) {
    long N = NCELL(p);
    int dims[3] = {p->nx, p->ny, p->nz};
    for (int f = 0; f < 6; f++) {
        if (p->bc_f[f] != 5) continue;
        int ax = f / 2, side = f & 1;
        int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
        for (int u1 = 0; u1 < dims[a1]; u1++)
            for (int u2 = 0; u2 < dims[a2]; u2++) {
                int cb[3], cn[3];
                cb[ax] = side ? dims[ax] - 1 : 0;
                cb[a1] = u1;
                cb[a2] = u2;
                for (int k = 0; k < 3; k++) cn[k] = cb[k];
                cn[ax] += side ? -1 : 1;
                long b = lin(p, cb[0], cb[1], cb[2]), n = lin(p, cn[0], cn[1], cn[2]);
                if (s->flag[b] == 3 || s->flag[n] == 3) continue;
                double rho = 0, j[3] = {0, 0, 0};
                for (int i = 0; i < q; i++) {
                    double v = post[i * N + n];
                    rho += v;
                    for (int d = 0; d < 3; d++) j[d] += E27[i][d] * v;
                }
                double un[3] = {j[0] / rho, j[1] / rho, j[2] / rho}, ub[3];
                wall_velocity(p, f, cb[0], cb[1], cb[2], ub);
                double eb[27], en[27];
                if (p->flow_incomp) {
                    orc_equilibrium_he(q, rho, p->rho0, ub, eb);
                    orc_equilibrium_he(q, rho, p->rho0, un, en);
                } else {
                    orc_equilibrium(q, rho, ub, eb);
                    orc_equilibrium(q, rho, un, en);
                }
#if ORC_MUTANT == 8
                for (int i = 0; i < q; i++) post[i * N + b] = post[i * N + n] - eb[i] + en[i];
#else
                for (int i = 0; i < q; i++) post[i * N + b] = post[i * N + n] + eb[i] - en[i];
#endif
            }
    }
}

/* Solute collision (mkSolutalField::propagate + relaxation) of every non-wall cell into post
 * (7 x N): dc/dt + div(c u) = div(D q grad U), U = c/A(phi), written as
 * dc/dt + div(c (u + V)) = div(D_c grad c) with D_c = D q / A and V = D q grad A / A^2.
 * phi_f / phi_prev: the phase field that drives the solute (NULL: liquid everywhere, no phase
 * terms); cpop: the 7 x N populations; Cm, Um: the cell values of c and U; uadv: 3 x N
 * advection velocity or NULL; per: the periodic axes of the phase-field gradient. */
static void solute_collide(const orc_params* p, const int8_t* flag, const double* phi_f,
                           const double* phi_prev, const double* cpop, const double* Cm,
                           const double* Um, const double* uadv, const int* per,
                           double* post) {
    long N = NCELL(p);
    lat_t L = make_lat(7);
    lat_t L19 = make_lat(19);
#pragma omp parallel for collapse(2) schedule(static)
    for (int X = 0; X < p->nx; X++)
        for (int Y = 0; Y < p->ny; Y++)
            for (int Z = 0; Z < p->nz; Z++) {
                long x = lin(p, X, Y, Z);
                if (flag[x] == 3) continue;
                double phi = phi_f ? phi_f[x] : -1.0;
                double gphi[3] = {0, 0, 0};
                if (phi_f) {
                    for (int i = 1; i < 19; i++) {
                        int XX = clampwrap(X + E27[i][0], p->nx, per[0]);
                        int YY = clampwrap(Y + E27[i][1], p->ny, per[1]);
                        int ZZ = clampwrap(Z + E27[i][2], p->nz, per[2]);
                        double v = 3.0 * L19.w[i] * phi_f[lin(p, XX, YY, ZZ)];
                        for (int a = 0; a < 3; a++) gphi[a] += E27[i][a] * v;
                    }
                }
                double q = 0.5 * (1.0 - phi);
                if (p->D_l > 0) q += (p->D_s / p->D_l) * 0.5 * (1.0 + phi);
                if (q < p->q_min) q = p->q_min;
                double A = 0.5 * ((1.0 + p->k_part) - (1.0 - p->k_part) * phi);
                double Dc = p->D_l * q / A;
                double om = 1.0 / (0.5 + Dc / L.cs2);
                double vfac = -0.5 * (1.0 - p->k_part) * p->D_l * q / (A * A);
                double dphi = phi_f ? (phi_f[x] - phi_prev[x]) : 0.0;
                double C = Cm[x];
                const double zero3[3] = {0, 0, 0};
                const double* u = uadv ? uadv + 3 * x : zero3;
                /* equilibrium flux c (u + V) + j_at, the anti-trapping current (Karma 2001)
                 * j_at = -a_t W (1 + (1-k) U) dphi/dt n, n = grad phi / |grad phi|, which
                 * moves solute from the growing solid to the liquid (mkSolutalField::ATcurrent,
                 * MaterKit/inc/mkSolutalField.h) */
                double j[3];
                for (int a = 0; a < 3; a++) j[a] = C * (u[a] + vfac * gphi[a]);
                double gg = gphi[0] * gphi[0] + gphi[1] * gphi[1] + gphi[2] * gphi[2];
                double atw = p->at_coef * sqrt(p->pf_wdx2 > 0 ? p->pf_wdx2 : 0.0);
                if (phi_f && atw != 0.0 && gg > 1e-20) {
                    double jat = -atw * (1.0 + (1.0 - p->k_part) * Um[x]) * dphi / sqrt(gg);
                    for (int a = 0; a < 3; a++) j[a] += jat * gphi[a];
                }
                double cs2_adv_C = L.cs2, latC = p->latent_C;
#if ORC_MUTANT == 6
                cs2_adv_C = 1.0 / 3.0;
#elif ORC_MUTANT == 7
                latC = -latC;
#endif
                for (int i = 0; i < 7; i++) {
                    double cj = E27[i][0] * j[0] + E27[i][1] * j[1] + E27[i][2] * j[2];
                    double ceq = L.w[i] * (C + cj / cs2_adv_C);
                    double ci = cpop[i * N + x];
                    post[i * N + x] = ci - om * (ci - ceq) + L.w[i] * latC * dphi;
                }
            }
}

/* Streaming of the solute lattice; its Dirichlet faces fix U: c_w = A(phi) U_w. */
static void solute_stream(const orc_params* p, const orc_state* s, const double* phi_f,
                          const double* post, double* c, double* fresh) {
    long N = NCELL(p);
    double* Ac = (double*)malloc((size_t)N * sizeof(double));
    for (long x = 0; x < N; x++) {
        double ph = phi_f ? phi_f[x] : -1.0;
        Ac[x] = 0.5 * ((1.0 + p->k_part) - (1.0 - p->k_part) * ph);
    }
    stream(p, s, 7, p->bc_c, 1, p->C_face, NULL, Ac, NULL, post, c, fresh);
    free(Ac);
    memcpy(c, fresh, (size_t)7 * N * sizeof(double));
}

/* ------------------------------------------------------------------ one step */
static void one_step(const orc_params* p, orc_state* s, int mask, int op) {
    long N = NCELL(p);
    int adv_f = (mask & 1) && p->flow_q;
    int adv_g = (mask & 2) && p->phase_q;
    int adv_h = (mask & 4) && p->thermal;
    int adv_c = (mask & 8) && p->solute;

    if (adv_g) {
        int qg = p->phase_q;
#pragma omp parallel for schedule(static)
        for (long x = 0; x < N; x++) {
            s->phi_prev[x] = s->phi[x];
            if (s->flag[x] == 3) continue; /* wall cells keep their phase value */
            double ph = 0;
            for (int i = 0; i < qg; i++) ph += s->g[i * N + x];
            s->phi[x] = ph;
            s->flag[x] = classify(p, s->flag[x], ph);
        }
    }

    macro_t m;
    m.rho = malloc(N * sizeof(double));
    m.u = malloc(3 * N * sizeof(double));
    m.T = malloc(N * sizeof(double));
    m.C = malloc(N * sizeof(double));
    m.U = malloc(N * sizeof(double));
    m.ns = malloc(N * sizeof(double));
    compute_macro(p, s, &m);

    double* post = NULL;
    size_t maxq = 27;
    post = malloc(maxq * N * sizeof(double));
    double* fresh = malloc(maxq * N * sizeof(double));

    /* flow */
    if (adv_f) {
        int q = p->flow_q;
#pragma omp parallel for schedule(static)
        for (long x = 0; x < N; x++) {
            double fin[27], fo[27], F[3];
            if (s->flag[x] == 3) continue;
            for (int i = 0; i < q; i++) fin[i] = s->f[i * N + x];
            body_force(p, m.T[x], F);
            orc_collide_cell(p, op, fin, F, fo, m.ns[x]);
            for (int i = 0; i < q; i++) post[i * N + x] = fo[i];
        }
        nee_faces(p, s, q, post);
        double* lf = (double*)malloc((size_t)N * sizeof(double));
        for (long x = 0; x < N; x++) lf[x] = 1.0 - m.ns[x];
        stream(p, s, q, p->bc_f, 0, NULL, lf, NULL, m.u, post, s->f, fresh);
        free(lf);
        memcpy(s->f, fresh, (size_t)q * N * sizeof(double));
    }

    /* advection velocity seen by the scalar lattices */
    double* uadv = malloc(3 * N * sizeof(double));
    for (long x = 0; x < N; x++)
        for (int a = 0; a < 3; a++) uadv[3 * x + a] = (1.0 - m.ns[x]) * m.u[3 * x + a];

    /* phase field */
    if (adv_g) {
        int q = p->phase_q;
        lat_t L = make_lat(q);
        double omg = 1.0 / (0.5 + p->pf_dt * p->pf_wdx2 / L.cs2);
        int per[3] = {p->bc_g[0] == 0, p->bc_g[2] == 0, p->bc_g[4] == 0};
#pragma omp parallel for collapse(2) schedule(static)
        for (int x = 0; x < p->nx; x++)
            for (int y = 0; y < p->ny; y++)
                for (int z = 0; z < p->nz; z++) {
                    long id = lin(p, x, y, z);
                    if (s->flag[id] == 3) continue;
                    double phi27[27];
                    for (int dx = -1; dx <= 1; dx++)
                        for (int dy = -1; dy <= 1; dy++)
                            for (int dz = -1; dz <= 1; dz++) {
                                int X = clampwrap(x + dx, p->nx, per[0]);
                                int Y = clampwrap(y + dy, p->ny, per[1]);
                                int Z = clampwrap(z + dz, p->nz, per[2]);
                                P27(dx, dy, dz) = s->phi[lin(p, X, Y, Z)];
                            }
                    double upf = 0;
                    if (p->pf_coupling == 1 && p->thermal) upf = m.T[id];
                    if (p->pf_coupling == 2 && p->solute) upf = m.U[id];
                    double S, a;
                    double nac[3] = {0, 0, 0}, zac = 0;
                    double ph = s->phi[id];
                    if (p->pf_model == 1) {
                        /* conservative Allen-Cahn (relaxation_AllenCahn): counter-term
                         * (1 - omega/2) w_i (c_i . n)(1 - phi^2) dx / (sqrt 2 W0) of Chai et
                         * al. (2018), n = s / (|s| + 1e-12), s = sum_i c_i w_i phi(x + c_i) on
                         * D3Q19 weights; growth source -lambda u (1 - phi^2)^2 */
                        lat_t L19 = make_lat(19);
                        double s3[3] = {0, 0, 0};
                        for (int k = 1; k < 19; k++)
                            for (int d = 0; d < 3; d++)
                                s3[d] += E27[k][d] * L19.w[k] * P27(E27[k][0], E27[k][1], E27[k][2]);
                        double sn = sqrt(s3[0] * s3[0] + s3[1] * s3[1] + s3[2] * s3[2]) + 1e-12;
                        for (int d = 0; d < 3; d++) nac[d] = s3[d] / sn;
                        double dd = 1.0 - ph * ph;
                        S = -p->lambda * upf * dd * dd * p->pf_dt;
                        zac = (1.0 - 0.5 * omg) * dd / (sqrt(2.0) * sqrt(p->pf_wdx2));
                    } else {
                        orc_phase_source(p, phi27, upf, s->grain ? s->grain[id] : 0, &S, &a);
                    }
                    const double* u = uadv + 3 * id;
                    for (int i = 0; i < q; i++) {
                        double cu = E27[i][0] * u[0] + E27[i][1] * u[1] + E27[i][2] * u[2];
                        double cn = E27[i][0] * nac[0] + E27[i][1] * nac[1] + E27[i][2] * nac[2];
                        /* the advective flux is (phi + 1) u: the liquid (phi = -1) is left
                         * unchanged by a velocity field with div u != 0 */
                        double geq = L.w[i] * (ph + (ph + 1.0) * cu / L.cs2);
                        double gi = s->g[i * N + id];
                        post[i * N + id] = gi - omg * (gi - geq) + L.w[i] * (S + zac * cn);
                    }
                }
        stream(p, s, q, p->bc_g, 1, NULL, NULL, NULL, NULL, post, s->g, fresh);
        memcpy(s->g, fresh, (size_t)q * N * sizeof(double));
    }

    /* thermal */
    if (adv_h) {
        lat_t L = make_lat(7);
#pragma omp parallel for schedule(static)
        for (long x = 0; x < N; x++) {
            if (s->flag[x] == 3) continue;
            double phi = p->phase_q ? s->phi[x] : -1.0;
            double fs = p->phase_q ? solid_fraction(phi) : 0.0;
            double alpha = p->alpha_l * (1.0 - fs) + p->alpha_s * fs;
            if (p->has_kappa && s->kappa) alpha *= s->kappa[x];
            double om = 1.0 / (0.5 + alpha / L.cs2);
            double dphi = p->phase_q ? (s->phi[x] - s->phi_prev[x]) : 0.0;
            double T = m.T[x];
            const double* u = uadv + 3 * x;
            /* heq = w T (1 + c.u / cs^2) with the D3Q7 cs^2 = 1/4 (sum_i w_i c_i c_i = I/4),
             * whose first moment is T u; latent source w_i L dphi, L = latent_T
             * (mkThermalField::relaxation(flow), MaterKit/inc/mkThermalField.h:33-36;
             * mkNode::temperature_collision, MaterKit/inc/mkNode.h:156) */
            double cs2_adv_T = L.cs2, latT = p->latent_T;
#if ORC_MUTANT == 1
            cs2_adv_T = 1.0 / 3.0;
#elif ORC_MUTANT == 2
            latT = -latT;
#elif ORC_MUTANT == 3
            latT = 2.0 * latT;
#endif
            for (int i = 0; i < 7; i++) {
                double cu = E27[i][0] * u[0] + E27[i][1] * u[1] + E27[i][2] * u[2];
                double heq = L.w[i] * T * (1.0 + cu / cs2_adv_T);
                double hi = s->h[i * N + x];
                post[i * N + x] = hi - om * (hi - heq) + L.w[i] * latT * dphi;
            }
        }
        stream(p, s, 7, p->bc_h, 1, p->T_face, NULL, NULL, NULL, post, s->h, fresh);
        memcpy(s->h, fresh, (size_t)7 * N * sizeof(double));
    }

    if (adv_c) {
        int per[3] = {p->bc_g[0] == 0, p->bc_g[2] == 0, p->bc_g[4] == 0};
        const double* ph = p->phase_q ? s->phi : NULL;
        solute_collide(p, s->flag, ph, s->phi_prev, s->c, m.C, m.U, uadv, per, post);
        solute_stream(p, s, ph, post, s->c, fresh);
    }

    free(uadv);
    free(post);
    free(fresh);
    free(m.rho); free(m.u); free(m.T); free(m.C); free(m.U); free(m.ns);
}

int orc_step(const orc_params* p, orc_state* s, int mask, int op, int nsteps) {
    if (p->flow_q && p->flow_q != 19 && p->flow_q != 27) return -1;
    if (p->phase_q && p->phase_q != 7 && p->phase_q != 19) return -1;
    for (int f = 0; f < 6; f++) {
        if (p->bc_g[f] == 5 || p->bc_h[f] == 5 || p->bc_c[f] == 5) return -1;
        /* velocity faces share no edge and stand at least three cells apart */
        if (p->bc_f[f] != 5) continue;
        int dims[3] = {p->nx, p->ny, p->nz};
        if (dims[f / 2] < 3) return -1;
        for (int g = 0; g < 6; g++)
            if (g / 2 != f / 2 && p->bc_f[g] == 5) return -1;
    }
    if (p->flow_q == 19) build_mrt(19, &MRT19);
    if (p->flow_q == 27) build_mrt(27, &MRT27);
    for (int n = 0; n < nsteps; n++) one_step(p, s, mask, op);
    return 0;
}

/* ------------------------------------------------------------------ initial states */
void orc_init_equilibrium(const orc_params* p, orc_state* s, const double* rho,
                          const double* vel3, const double* T, const double* C,
                          const double* phi) {
    long N = NCELL(p);
    for (long x = 0; x < N; x++) {
        double ph = phi ? phi[x] : 0.0;
        double u[3] = {0, 0, 0};
        if (vel3)
            for (int a = 0; a < 3; a++) u[a] = vel3[3 * x + a];
        if (p->phase_q && phi) {
            lat_t L = make_lat(p->phase_q);
            for (int i = 0; i < L.q; i++) s->g[i * N + x] = L.w[i] * ph;
            s->phi[x] = ph;
            s->phi_prev[x] = ph;
            s->flag[x] = classify(p, s->flag[x], ph);
        }
        if (p->flow_q) {
            double feq[27];
            if (p->flow_incomp)
                orc_equilibrium_he(p->flow_q, rho ? rho[x] : p->rho0, p->rho0, u, feq);
            else
                orc_equilibrium(p->flow_q, rho ? rho[x] : p->rho0, u, feq);
            for (int i = 0; i < p->flow_q; i++) s->f[i * N + x] = feq[i];
        }
        lat_t L7 = make_lat(7);
        if (p->thermal) {
            double t = T ? T[x] : p->T_ref;
            for (int i = 0; i < 7; i++) s->h[i * N + x] = L7.w[i] * t;
        }
        if (p->solute) {
            double c = C ? C[x] : 0.0;
            for (int i = 0; i < 7; i++) s->c[i * N + x] = L7.w[i] * c;
        }
    }
}

void orc_init_seeds(const orc_params* p, orc_state* s, int n, const double* centers,
                    const double* radii) {
    /* interface half-width in cells: sqrt(2) W0 / dx  (W0/dx = sqrt(pf_wdx2)) */
    double width = sqrt(2.0) * sqrt(p->pf_wdx2);
    double ts = atanh(p->phi_s) * width;
    for (int x = 0; x < p->nx; x++)
        for (int y = 0; y < p->ny; y++)
            for (int z = 0; z < p->nz; z++) {
                long id = lin(p, x, y, z);
                double best = 1e300;
                int gbest = 0;
                for (int k = 0; k < n; k++) {
                    double dx = x - centers[3 * k], dy = y - centers[3 * k + 1],
                           dz = z - centers[3 * k + 2];
                    double d = sqrt(dx * dx + dy * dy + dz * dz) - radii[k];
                    if (d < best) {
                        best = d;
                        gbest = k + 1;
                    }
                }
                double ph = -tanh(best / width);
                s->phi[id] = ph;
                s->phi_prev[id] = ph;
                {
                    lat_t L = make_lat(p->phase_q ? p->phase_q : 19);
                    if (p->phase_q)
                        for (int i = 0; i < L.q; i++) s->g[i * NCELL(p) + id] = L.w[i] * ph;
                }
                if (s->grain) s->grain[id] = (int16_t)gbest;
                if (s->flag[id] != 3 && s->flag[id] >= 0 && p->pf_model == 0)
                    s->flag[id] = (best <= -ts) ? 2 : (best < ts ? 1 : 0);
            }
}

/* ------------------------------------------------------------------ multi-phase field
 * n order parameters phi_j in [-1, 1], each with its own D3Q19 population lattice: the gVec
 * lattices of mkFluidsUser (MaterKit/src/mkFluidsUser.cpp:68-299, phase_number of
 * MaterKit/inc/mkMultiphaseField.h:56), run by run_multiPhaseField (MaterKit/MaterKit.cpp:
 * 529-576).  One step, following the visible loops:
 *   1. phi_j = sum_i g_j,i and the combined phi = (n - 1) + sum_j phi_j (:147-172)
 *   2. per non-wall cell and order parameter:
 *        s_j = sum_i c_i w_i phi_j(x + c_i), n_j = s_j / (|s_j| + eps)       (:208-217)
 *        g_eq = w_i phi_j (1 + 3 c.u_j + 4.5 (c.u_j)^2 - 1.5 u_j^2): getEquPDFs with phi_j
 *               as the density and the order parameter's own velocity u_j          (:219-230)
 *        P_j = -2 kappa_j (phi_j/2 + 1/2) sum_{k != j} (phi_k/2 + 1/2)^2           (:236-256)
 *        Q_j = growth (phi_j - lambda T (1 - phi_j^2)) (1 - phi_j^2)   (growth 0 there, :266-271)
 *        Z_i = (1 - omega_g/2) w_i (c_i . n_j) (1 - phi_j^2) dx / (sqrt(2) W0): the
 *              conservative Allen-Cahn counter-term of getAddedPhasePDFs(omega_g, phi_j, n_j)
 *              in the form of Chai, Sun, Wang & Shi (2018) (:273-279)
 *        g* = g - omega_g (g - g_eq) + w_i dt (P_j + Q_j) + Z_i                     (:285-293)
 *      with omega_g = 1 / (dt/dx^2/cs^2 + 1/2) (:282)
 *   3. pull streaming with the face rules of the lattice (:68-134)
 * The combined lattice g of the template is not evolved: its density is overwritten by the
 * combined phi at every step (:171-172), so it feeds nothing.
 * Optional D3Q7 thermal lattice (run_multiPhaseField's mkThermalField calls, :566-567): BGK
 * diffusion with alpha, no advection, latent source w_i L dphi of the combined phi; T enters
 * Q_j.
 *
 * Karma-Rappel grains (model 1; mkMultiphaseField::relaxation_anisotropy_thermal,
 * relaxation_solute, relaxation_thermal with anisotropy / _facet / _cbc / _hcp per grain,
 * MaterKit/inc/mkMultiphaseField.h:31-45, whose bodies are not public): each order parameter
 * is the anisotropic phase field of the single-grain solver (orc_phase_source) with the
 * grain's own form, strength and rotation, driven by u = T (coupling 1) or U (coupling 2), and
 * the pairwise interaction joins the source on the same time scale tau0 a_s^2:
 *        g* = g - omega_g (g - g_eq) + w_i [S_j + dt P_j / a_j^2]
 * without the Allen-Cahn counter-term.  With one grain this is the single-grain phase field
 * exactly.  Optional D3Q7 solute lattice (solute 1): the single-grain solute equation
 * (solute_collide) driven by the combined phi clamped to [-1, 1] (grains that overlap take it
 * above 1, where A(phi) = ((1 + k) - (1 - k) phi) / 2 would vanish), U = c / A(phi). */
static orc_params mpf_geom(const orc_mpf_params* m, const int* bc) {
    orc_params p;
    memset(&p, 0, sizeof(p));
    p.nx = m->nx;
    p.ny = m->ny;
    p.nz = m->nz;
    for (int f = 0; f < 6; f++) p.bc_g[f] = bc[f];
    p.rho0 = 1.0;
    return p;
}

void orc_mpf_init(const orc_mpf_params* m, const double* phij, const double* T, const double* C,
                  double* g, double* phi, double* phi_prev, double* h, double* c) {
    long N = (long)m->nx * m->ny * m->nz;
    lat_t L7 = make_lat(7);
    for (long x = 0; x < N; x++) {
        double comb = m->n - 1.0;
        for (int j = 0; j < m->n; j++) {
            double ph = phij[j * N + x];
            const double* u = m->u[j];
            double feq[27];
            orc_equilibrium(19, ph, u, feq);
            for (int i = 0; i < 19; i++) g[((long)j * 19 + i) * N + x] = feq[i];
            phi[j * N + x] = phi_prev[j * N + x] = ph;
            comb += ph;
        }
        phi[(long)m->n * N + x] = phi_prev[(long)m->n * N + x] = comb;
        if (h) {
            double t = T ? T[x] : m->T_ref;
            for (int i = 0; i < 7; i++) h[i * N + x] = L7.w[i] * t;
        }
        if (c) {
            double cc = C ? C[x] : 0.0;
            for (int i = 0; i < 7; i++) c[i * N + x] = L7.w[i] * cc;
        }
    }
}

/* The single-grain phase-field parameters of grain j (form, strength, rotation). */
static orc_params mpf_grain(const orc_mpf_params* m, int j) {
    orc_params p = mpf_geom(m, m->bc);
    p.phase_q = 19;
    p.aniso_kind = m->aniso_kind[j];
    p.eps4 = m->eps4[j];
    p.aniso[0] = m->aniso[j][0];
    p.aniso[1] = m->aniso[j][1];
    p.ngrains = 1;
    memcpy(p.rot[0], m->rot[j], sizeof(p.rot[0]));
    p.lambda = m->lambda;
    p.pf_dt = m->dt;
    p.pf_wdx2 = m->wdx * m->wdx;
    return p;
}

/* The single-grain solute parameters of the solute lattice. */
static orc_params mpf_solute(const orc_mpf_params* m) {
    orc_params p = mpf_geom(m, m->bc);
    p.phase_q = 19;
    p.solute = 1;
    for (int f = 0; f < 6; f++) {
        p.bc_c[f] = m->bc_c[f];
        p.C_face[f] = m->C_face[f];
    }
    p.D_l = m->D_l;
    p.D_s = m->D_s;
    p.k_part = m->k_part;
    p.q_min = m->q_min;
    p.latent_C = m->latent_C;
    p.at_coef = m->at_coef;
    p.pf_wdx2 = m->wdx * m->wdx;
    return p;
}

int orc_mpf_step(const orc_mpf_params* m, double* g, double* phi, double* phi_prev, double* h,
                 double* c, int8_t* flag, int nsteps) {
    if (m->n < 1 || m->n > ORC_MPF_MAX) return -1;
    if (m->model != 0 && m->model != 1) return -1;
    if ((m->thermal != 0) != (h != NULL) || (m->solute != 0) != (c != NULL)) return -1;
    long N = (long)m->nx * m->ny * m->nz;
    int n = m->n;
    lat_t L = make_lat(19);
    lat_t L7 = make_lat(7);
    orc_params pg = mpf_geom(m, m->bc);
    orc_params ph7 = mpf_geom(m, m->bc_h);
    orc_params ps = mpf_solute(m);
    int8_t* fl = flag;
    int8_t* own = NULL;
    if (!fl) {
        own = calloc((size_t)N, 1);
        fl = own;
    }
    orc_state st;
    memset(&st, 0, sizeof(st));
    st.flag = fl;
    double omg = 1.0 / (0.5 + m->dt * m->wdx * m->wdx / L.cs2);
    double lam = 1.0 / (sqrt(2.0) * m->wdx);
    int per[3] = {m->bc[0] == 0, m->bc[2] == 0, m->bc[4] == 0};
    double* post = malloc((size_t)19 * N * sizeof(double));
    double* fresh = malloc((size_t)19 * N * sizeof(double));
    for (int step = 0; step < nsteps; step++) {
        /* 1. moments */
        for (long x = 0; x < N; x++) {
            for (int j = 0; j <= n; j++) phi_prev[j * N + x] = phi[j * N + x];
            if (fl[x] == 3) continue;
            double comb = n - 1.0;
            for (int j = 0; j < n; j++) {
                double s = 0;
                for (int i = 0; i < 19; i++) s += g[((long)j * 19 + i) * N + x];
                phi[j * N + x] = s;
                comb += s;
            }
            phi[(long)n * N + x] = comb;
        }
        /* thermal and solute moments at this level (T or U drives the grains) */
        double* Tc = NULL;
        if (h) {
            Tc = malloc((size_t)N * sizeof(double));
            for (long x = 0; x < N; x++) {
                double t = 0;
                for (int i = 0; i < 7; i++) t += h[i * N + x];
                Tc[x] = t;
            }
        }
        /* the solute sees the combined phi clamped to [-1, 1] (overlapping grains exceed 1) */
        double *Cc = NULL, *Uc = NULL, *ps_now = NULL, *ps_prev = NULL;
        if (c) {
            Cc = malloc((size_t)N * sizeof(double));
            Uc = malloc((size_t)N * sizeof(double));
            ps_now = malloc((size_t)N * sizeof(double));
            ps_prev = malloc((size_t)N * sizeof(double));
            for (long x = 0; x < N; x++) {
                ps_now[x] = fmin(1.0, fmax(-1.0, phi[(long)n * N + x]));
                ps_prev[x] = fmin(1.0, fmax(-1.0, phi_prev[(long)n * N + x]));
                double cs = 0;
                for (int i = 0; i < 7; i++) cs += c[i * N + x];
                double A = 0.5 * ((1.0 + m->k_part) - (1.0 - m->k_part) * ps_now[x]);
                Cc[x] = cs;
                Uc[x] = cs / A;
            }
        }
        /* 2. collision and 3. streaming, one order parameter at a time */
        for (int j = 0; j < n; j++) {
            const double* u = m->u[j];
            orc_params pj = mpf_grain(m, j);
#pragma omp parallel for collapse(2) schedule(static)
            for (int X = 0; X < m->nx; X++)
                for (int Y = 0; Y < m->ny; Y++)
                    for (int Z = 0; Z < m->nz; Z++) {
                        long x = lin(&pg, X, Y, Z);
                        if (fl[x] == 3) continue;
                        double pj_ = phi[j * N + x];
                        double P = 0;
                        for (int k = 0; k < n; k++)
                            if (k != j) {
                                double hk = 0.5 * phi[k * N + x] + 0.5;
                                P += hk * hk;
                            }
                        P *= -2.0 * m->kappa[j] * (0.5 * pj_ + 0.5);
                        double feq[27];
                        orc_equilibrium(19, pj_, u, feq);
                        if (m->model == 1) {
                            double phi27[27];
                            for (int dx = -1; dx <= 1; dx++)
                                for (int dy = -1; dy <= 1; dy++)
                                    for (int dz = -1; dz <= 1; dz++) {
                                        int XX = clampwrap(X + dx, m->nx, per[0]);
                                        int YY = clampwrap(Y + dy, m->ny, per[1]);
                                        int ZZ = clampwrap(Z + dz, m->nz, per[2]);
                                        P27(dx, dy, dz) = phi[j * N + lin(&pg, XX, YY, ZZ)];
                                    }
                            double upf = 0;
                            if (m->coupling == 1 && Tc) upf = Tc[x];
                            if (m->coupling == 2 && Uc) upf = Uc[x];
                            double S, a;
                            orc_phase_source(&pj, phi27, upf, 1, &S, &a);
                            double src = S + m->dt * P / (a * a);
                            for (int i = 0; i < 19; i++) {
                                double gi = g[((long)j * 19 + i) * N + x];
                                post[i * N + x] = gi - omg * (gi - feq[i]) + L.w[i] * src;
                            }
                            continue;
                        }
                        double s3[3] = {0, 0, 0};
                        for (int i = 1; i < 19; i++) {
                            int XX = clampwrap(X + E27[i][0], m->nx, per[0]);
                            int YY = clampwrap(Y + E27[i][1], m->ny, per[1]);
                            int ZZ = clampwrap(Z + E27[i][2], m->nz, per[2]);
                            double v = L.w[i] * phi[j * N + lin(&pg, XX, YY, ZZ)];
                            for (int a = 0; a < 3; a++) s3[a] += E27[i][a] * v;
                        }
                        double sn = sqrt(s3[0] * s3[0] + s3[1] * s3[1] + s3[2] * s3[2]) + m->eps;
                        double nv[3] = {s3[0] / sn, s3[1] / sn, s3[2] / sn};
                        double d = 1.0 - pj_ * pj_;
                        double T = Tc ? Tc[x] : m->T_ref;
                        double Q = m->growth * (pj_ - m->lambda * T * d) * d;
                        for (int i = 0; i < 19; i++) {
                            double cn = E27[i][0] * nv[0] + E27[i][1] * nv[1] + E27[i][2] * nv[2];
                            double Zc = (1.0 - 0.5 * omg) * L.w[i] * cn * d * lam;
                            double gi = g[((long)j * 19 + i) * N + x];
                            post[i * N + x] = gi - omg * (gi - feq[i]) +
                                              L.w[i] * m->dt * (P + Q) + Zc;
                        }
                    }
            double* gj = g + (long)j * 19 * N;
            stream(&pg, &st, 19, m->bc, 1, m->phi_face, NULL, NULL, NULL, post, gj, fresh);
            memcpy(gj, fresh, (size_t)19 * N * sizeof(double));
        }
        /* thermal */
        if (h) {
            for (long x = 0; x < N; x++) {
                if (fl[x] == 3) continue;
                double om = 1.0 / (0.5 + m->alpha / L7.cs2);
                double dphi = phi[(long)n * N + x] - phi_prev[(long)n * N + x];
                for (int i = 0; i < 7; i++) {
                    double hi = h[i * N + x];
                    double heq = L7.w[i] * Tc[x];
                    post[i * N + x] = hi - om * (hi - heq) + L7.w[i] * m->latent_T * dphi;
                }
            }
            stream(&ph7, &st, 7, m->bc_h, 1, m->T_face, NULL, NULL, NULL, post, h, fresh);
            memcpy(h, fresh, (size_t)7 * N * sizeof(double));
            free(Tc);
        }
        /* solute, driven by the clamped combined phi */
        if (c) {
            solute_collide(&ps, fl, ps_now, ps_prev, c, Cc, Uc, NULL, per, post);
            solute_stream(&ps, &st, ps_now, post, c, fresh);
            free(Cc);
            free(Uc);
            free(ps_now);
            free(ps_prev);
        }
    }
    free(post);
    free(fresh);
    free(own);
    return 0;
}

/* ------------------------------------------------------------------ finite-volume solvers
 * Explicit cell-centred finite-volume forms of the phase field (evolution_anisotropy,
 * MaterKit/inc/mkPhaseField.h:51, "on FVM") and of the dilute-alloy solute with the
 * anti-trapping current (evolution_ATC1 / evolution_ATC2, MaterKit/inc/mkSolutalField.h:44-46;
 * run_solutal_only, MaterKit/MaterKit.cpp:689-741).  The bodies are in the closed DLL; the
 * forms are the published Karma-Rappel (1998) and Karma (2001) / Echebarria et al. (2004)
 * equations on the same stencils as the lattice solver:
 *   phase:  phi' = phi + dt/a^2 (W^2/dx^2 div J + Q),  J on the six faces as in
 *           orc_phase_source (so phi' - phi = S_LB + dt W^2/dx^2 lap7 phi, S_LB its source)
 *   solute: c = A(phi) U,  A = ((1+k) - (1-k) phi)/2,
 *           c' = c - sum_faces J_out + (phi' - phi)/2,
 *           J_out = -D q_f (U_nb - U) + s j_at,a   (face with outward side s on axis a)
 *           q_f = (q(phi) + q(phi_nb))/2, q = (1-phi)/2 + (D_s/D_l)(1+phi)/2 >= q_min
 *           j_at,a = -a_t W/dx (1 + (1-k) U_f) dphi_f g_a/|g|, U_f and dphi_f face means,
 *           g the face gradient of phi (normal difference, transverse central means)
 *   faces:  periodic; bounce = zero flux (phi replicated at the edge); Dirichlet (solute)
 *           fixes U at the face: J_out = -2 D q (U_w - U)
 * c - phi/2 is conserved exactly with periodic and zero-flux faces. */
static double fvm_q(const orc_params* p, double phi) {
    double q = 0.5 * (1.0 - phi);
    if (p->D_l > 0) q += (p->D_s / p->D_l) * 0.5 * (1.0 + phi);
    return q < p->q_min ? p->q_min : q;
}
static double fvm_A(const orc_params* p, double phi) {
    return 0.5 * ((1.0 + p->k_part) - (1.0 - p->k_part) * phi);
}

/* The phase pass of one finite-volume step: phin = phi + S + dt W^2/dx^2 lap7 phi with the
 * lattice solver's source S driven by u[x] (evolve_phase 0: phin = phi). */
static void fvm_phase_pass(const orc_params* p, const double* phi, const double* u, double* phin,
                           int evolve_phase) {
    int per[3] = {p->bc_g[0] == 0, p->bc_g[2] == 0, p->bc_g[4] == 0};
#pragma omp parallel for collapse(2) schedule(static)
    for (int X = 0; X < p->nx; X++)
        for (int Y = 0; Y < p->ny; Y++)
            for (int Z = 0; Z < p->nz; Z++) {
                long id = lin(p, X, Y, Z);
                if (!evolve_phase) {
                    phin[id] = phi[id];
                    continue;
                }
                double phi27[27];
                for (int dx = -1; dx <= 1; dx++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dz = -1; dz <= 1; dz++) {
                            int XX = clampwrap(X + dx, p->nx, per[0]);
                            int YY = clampwrap(Y + dy, p->ny, per[1]);
                            int ZZ = clampwrap(Z + dz, p->nz, per[2]);
                            P27(dx, dy, dz) = phi[lin(p, XX, YY, ZZ)];
                        }
                double S, a;
                orc_phase_source(p, phi27, u[id], 0, &S, &a);
                double lap = P27(1, 0, 0) + P27(-1, 0, 0) + P27(0, 1, 0) + P27(0, -1, 0) +
                             P27(0, 0, 1) + P27(0, 0, -1) - 6.0 * P27(0, 0, 0);
                phin[id] = phi[id] + S + p->pf_dt * p->pf_wdx2 * lap;
            }
}

int orc_fvm_step(const orc_params* p, double* phi, double* c, int nsteps, int evolve_phase) {
    long N = NCELL(p);
    int per[3] = {p->bc_g[0] == 0, p->bc_g[2] == 0, p->bc_g[4] == 0};
    int perc[3] = {p->bc_c[0] == 0, p->bc_c[2] == 0, p->bc_c[4] == 0};
    double* phin = malloc((size_t)N * sizeof(double));
    double* cn = malloc((size_t)N * sizeof(double));
    double* U = malloc((size_t)N * sizeof(double));
    double atw = p->at_coef * sqrt(p->pf_wdx2 > 0 ? p->pf_wdx2 : 0.0);
    int dims[3] = {p->nx, p->ny, p->nz};
    for (int step = 0; step < nsteps; step++) {
        for (long x = 0; x < N; x++) U[x] = c[x] / fvm_A(p, phi[x]);
        fvm_phase_pass(p, phi, U, phin, evolve_phase);
        /* solute */
#pragma omp parallel for collapse(2) schedule(static)
        for (int X = 0; X < p->nx; X++)
            for (int Y = 0; Y < p->ny; Y++)
                for (int Z = 0; Z < p->nz; Z++) {
                    long id = lin(p, X, Y, Z);
                    int P3[3] = {X, Y, Z};
                    double out = 0; /* sum of outward fluxes */
                    for (int ax = 0; ax < 3; ax++)
                        for (int s = 1; s >= -1; s -= 2) {
                            int Q3[3] = {X, Y, Z};
                            Q3[ax] += s;
                            int face = 2 * ax + (s > 0 ? 1 : 0);
                            if (Q3[ax] < 0 || Q3[ax] >= dims[ax]) {
                                if (perc[ax]) {
                                    Q3[ax] = (Q3[ax] + dims[ax]) % dims[ax];
                                } else {
                                    if (p->bc_c[face] == 2) /* Dirichlet: U fixed at the face */
                                        out += -2.0 * p->D_l * fvm_q(p, phi[id]) *
                                               (p->C_face[face] - U[id]);
                                    continue;
                                }
                            }
                            long nb = lin(p, Q3[0], Q3[1], Q3[2]);
                            double qf = 0.5 * (fvm_q(p, phi[id]) + fvm_q(p, phi[nb]));
                            double J = -p->D_l * qf * (U[nb] - U[id]);
                            if (atw != 0.0) {
                                /* face gradient of phi^n: normal difference, transverse means */
                                double g[3];
                                for (int b = 0; b < 3; b++) {
                                    if (b == ax) {
                                        g[b] = s * (phi[nb] - phi[id]);
                                        continue;
                                    }
                                    double v = 0;
                                    for (int side = 0; side < 2; side++) {
                                        int* C3 = side ? Q3 : P3;
                                        int A3[3] = {C3[0], C3[1], C3[2]}, B3[3] = {C3[0], C3[1], C3[2]};
                                        A3[b] = clampwrap(C3[b] + 1, dims[b], per[b]);
                                        B3[b] = clampwrap(C3[b] - 1, dims[b], per[b]);
                                        v += phi[lin(p, A3[0], A3[1], A3[2])] - phi[lin(p, B3[0], B3[1], B3[2])];
                                    }
                                    g[b] = 0.25 * v;
                                }
                                double gg = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
                                if (gg > 1e-20) {
                                    double Uf = 0.5 * (U[id] + U[nb]);
                                    double df = 0.5 * ((phin[id] - phi[id]) + (phin[nb] - phi[nb]));
                                    double ja = -atw * (1.0 + (1.0 - p->k_part) * Uf) * df * g[ax] / sqrt(gg);
                                    J += s * ja;
                                }
                            }
                            out += J;
                        }
                    cn[id] = c[id] - out + 0.5 * (phin[id] - phi[id]);
                }
        memcpy(phi, phin, (size_t)N * sizeof(double));
        memcpy(c, cn, (size_t)N * sizeof(double));
    }
    free(phin);
    free(cn);
    free(U);
    return 0;
}

/* Finite-volume phase field driven by the temperature (mkSolutalField::evolution_thermal,
 * MaterKit/inc/mkSolutalField.h:47-48: "potential is temperature"): the phase pass of
 * orc_fvm_step with u = T, then the heat equation with the latent source on the same faces,
 *   T' = T - sum_faces J_out + L_T (phi' - phi),   J_out = -alpha_f (T_nb - T),
 *   alpha_f = (alpha(phi) + alpha(phi_nb))/2,  alpha = alpha_l (1 - f_s) + alpha_s f_s,
 *   f_s = (1 + phi)/2 clamped to [0, 1]  (the thermal lattice's alpha(phi));
 *   faces bc_h: periodic; bounce = zero flux; Dirichlet fixes T at the face:
 *   J_out = -2 alpha(phi) (T_w - T).
 * T - L_T phi is conserved exactly with periodic and zero-flux faces. */
int orc_fvm_thermal_step(const orc_params* p, double* phi, double* T, int nsteps,
                         int evolve_phase) {
    long N = NCELL(p);
    int pert[3] = {p->bc_h[0] == 0, p->bc_h[2] == 0, p->bc_h[4] == 0};
    double* phin = malloc((size_t)N * sizeof(double));
    double* Tn = malloc((size_t)N * sizeof(double));
    int dims[3] = {p->nx, p->ny, p->nz};
    for (int step = 0; step < nsteps; step++) {
        fvm_phase_pass(p, phi, T, phin, evolve_phase);
#pragma omp parallel for collapse(2) schedule(static)
        for (int X = 0; X < p->nx; X++)
            for (int Y = 0; Y < p->ny; Y++)
                for (int Z = 0; Z < p->nz; Z++) {
                    long id = lin(p, X, Y, Z);
                    double a0;
                    {
                        double fs = 0.5 * (1.0 + phi[id]);
                        fs = fs < 0 ? 0 : (fs > 1 ? 1 : fs);
                        a0 = p->alpha_l * (1.0 - fs) + p->alpha_s * fs;
                    }
                    double out = 0;
                    for (int ax = 0; ax < 3; ax++)
                        for (int s = 1; s >= -1; s -= 2) {
                            int Q3[3] = {X, Y, Z};
                            Q3[ax] += s;
                            int face = 2 * ax + (s > 0 ? 1 : 0);
                            if (Q3[ax] < 0 || Q3[ax] >= dims[ax]) {
                                if (pert[ax]) {
                                    Q3[ax] = (Q3[ax] + dims[ax]) % dims[ax];
                                } else {
                                    if (p->bc_h[face] == 2) /* Dirichlet: T fixed at the face */
                                        out += -2.0 * a0 * (p->T_face[face] - T[id]);
                                    continue;
                                }
                            }
                            long nb = lin(p, Q3[0], Q3[1], Q3[2]);
                            double fs = 0.5 * (1.0 + phi[nb]);
                            fs = fs < 0 ? 0 : (fs > 1 ? 1 : fs);
                            double an = p->alpha_l * (1.0 - fs) + p->alpha_s * fs;
                            out += -0.5 * (a0 + an) * (T[nb] - T[id]);
                        }
                    Tn[id] = T[id] - out + p->latent_T * (phin[id] - phi[id]);
                }
        memcpy(phi, phin, (size_t)N * sizeof(double));
        memcpy(T, Tn, (size_t)N * sizeof(double));
    }
    free(phin);
    free(Tn);
    return 0;
}

/* ------------------------------------------------------------------ two-phase flow
 * Phase-field two-phase flow of mkFluidsMultiphase (MaterKit/inc/mkFluidsMultiphase.h:20-99;
 * run_flow_multiphase, MaterKit/MaterKit.cpp:290-331): its parameters are those of Liang et al.
 * (PRE 89, 053320, 2014) -- rho_l, rho_g, sigma, W_0, beta = 12 sigma / W, kappa =
 * 1.5 sigma W, eta "controls the mobility" (mkFluidsMultiphase.h:81-97); the bodies are in the
 * closed DLL, so the published model is restated (BGK collisions, per-cell tau_f):
 *   order parameter phi in [0, 1] (1: phase of rho_l), rho = rho_g + phi (rho_l - rho_g)
 *   mu = 4 beta phi (phi - 1)(phi - 1/2) - kappa lap phi
 *   Cahn-Hilliard lattice h (D3Q19): h0eq = phi + (w0 - 1) eta mu, hieq = w_i eta mu +
 *     w_i phi c_i.u / cs^2;  source (1 - 1/(2 tau_h)) w_i c_i.[(phi u)^n - (phi u)^(n-1)] / cs^2;
 *     M = eta cs^2 (tau_h - 1/2)
 *   flow lattice f (D3Q19): f0eq = (w0 - 1) p / cs^2 + rho s_0(u), fieq = w_i p / cs^2 +
 *     rho s_i(u), s_i = w_i (c.u/cs^2 + (c.u)^2/(2 cs^4) - u.u/(2 cs^2));
 *     rho u = sum c f + F/2,  p = cs^2/(1 - w0) [sum_{i>0} f_i + u.grad rho/2 + rho s_0(u)],
 *     F = mu grad phi + rho g + force;  forcing (1 - 1/(2 tau_f)) w_i [c_i.F +
 *     (c_i.u)(c_i.grad rho)]/cs^2, whose zeroth moment u.grad rho keeps div u = 0 where rho
 *     varies (the u.grad rho/2 of p is its half-step part);  nu (mp_nu) = nu_g + phi (nu_l - nu_g),
 *     or from mu = rho nu linear or harmonic in phi,
 *     tau_f = nu/cs^2 + 1/2
 *   gradients and Laplacian: isotropic D3Q19 forms, grad a = sum w_i c_i a(x+c_i)/cs^2,
 *     lap a = 2 sum w_i (a(x+c_i) - a(x))/cs^2; non-periodic faces replicate the edge value
 *   faces: periodic, or bounce (no-slip wall for f, zero flux for h).
 *   wetting walls (initialize_wetting, mkFluidsMultiphase.h:35, theta :95): surface energy
 *     -sigma cos(theta) phi^2 (3 - 2 phi) (Young's law with theta in the phi = 1 fluid);
 *     minimising the free energy gives kappa n.grad phi = 6 sigma cos(theta) phi (1 - phi),
 *     n the outward normal, so with kappa = 1.5 sigma W the ghost value one cell across the
 *     wall is phi_e + (4/W) cos(theta) phi_e (1 - phi_e), once per wetting face crossed.
 * One step: phi = sum h; mu; collide both lattices; stream. */
static void mp_grad_lap(const orc_mp_params* m, const double* a, int X, int Y, int Z,
                        double* grad, double* lap) {
    orc_params g = {0};
    g.nx = m->nx; g.ny = m->ny; g.nz = m->nz;
    int per[3] = {m->bc[0] == 0, m->bc[2] == 0, m->bc[4] == 0};
    int n[3] = {m->nx, m->ny, m->nz}, P[3] = {X, Y, Z};
    lat_t L = make_lat(19);
    double a0 = a[lin(&g, X, Y, Z)];
    double kw = m->wet_faces ? 4.0 / m->width * cos(m->contact_angle * (3.14159265358979323846 / 180.0)) : 0.0;
    double gr[3] = {0, 0, 0}, lp = 0;
    for (int i = 1; i < 19; i++) {
        int XX = clampwrap(X + E27[i][0], m->nx, per[0]);
        int YY = clampwrap(Y + E27[i][1], m->ny, per[1]);
        int ZZ = clampwrap(Z + E27[i][2], m->nz, per[2]);
        double v = a[lin(&g, XX, YY, ZZ)];
        int crossed = 0;
        for (int d = 0; d < 3; d++) {
            int q = P[d] + E27[i][d];
            if (q < 0 && (m->wet_faces >> (2 * d) & 1)) crossed++;
            if (q >= n[d] && (m->wet_faces >> (2 * d + 1) & 1)) crossed++;
        }
        if (crossed) v += crossed * kw * v * (1.0 - v);
        for (int d = 0; d < 3; d++) gr[d] += L.w[i] * E27[i][d] * v / L.cs2;
        lp += 2.0 * L.w[i] * (v - a0) / L.cs2;
    }
    if (grad) for (int d = 0; d < 3; d++) grad[d] = gr[d];
    if (lap) *lap = lp;
}

static void mp_phi_mu(const orc_mp_params* m, const double* h, double* phi, double* mu) {
    long N = (long)m->nx * m->ny * m->nz;
    for (long x = 0; x < N; x++) {
        double s = 0;
        for (int i = 0; i < 19; i++) s += h[i * N + x];
        phi[x] = s;
    }
    double beta = 12.0 * m->sigma / m->width, kap = 1.5 * m->sigma * m->width;
    orc_params g = {0};
    g.nx = m->nx; g.ny = m->ny; g.nz = m->nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int X = 0; X < m->nx; X++)
        for (int Y = 0; Y < m->ny; Y++)
            for (int Z = 0; Z < m->nz; Z++) {
                long id = lin(&g, X, Y, Z);
                double lp;
                mp_grad_lap(m, phi, X, Y, Z, NULL, &lp);
                double ph = phi[id];
                mu[id] = 4.0 * beta * ph * (ph - 1.0) * (ph - 0.5) - kap * lp;
            }
}

/* Kinematic viscosity at a clamped phi: linear in nu (visc_model 0), linear in the dynamic
 * viscosity mu = rho nu (1), or harmonic in mu, 1/mu = phi/mu_l + (1 - phi)/mu_g (2). */
static double mp_nu(const orc_mp_params* m, double pc) {
    if (m->visc_model == 0) return m->nu_g + pc * (m->nu_l - m->nu_g);
    double rho = m->rho_g + pc * (m->rho_l - m->rho_g);
    double ml = m->rho_l * m->nu_l, mg = m->rho_g * m->nu_g;
    double mu = m->visc_model == 1 ? mg + pc * (ml - mg) : 1.0 / (pc / ml + (1.0 - pc) / mg);
    return mu / rho;
}

/* Macroscopic fields of a cell from its pre-collision f (full values), phi and mu. */
static void mp_macro_cell(const orc_mp_params* m, const double* f, double ph, double muc,
                          const double* gphi, double* rho_o, double* u, double* p_o, double* F) {
    lat_t L = make_lat(19);
    double rho = m->rho_g + ph * (m->rho_l - m->rho_g);
    double grho[3];
    for (int d = 0; d < 3; d++) {
        grho[d] = (m->rho_l - m->rho_g) * gphi[d];
        F[d] = muc * gphi[d] + rho * m->gravity[d] + m->force[d];
    }
    double j[3] = {0, 0, 0}, s1 = 0;
    for (int i = 1; i < 19; i++) {
        s1 += f[i];
        for (int d = 0; d < 3; d++) j[d] += E27[i][d] * f[i];
    }
    for (int d = 0; d < 3; d++) u[d] = (j[d] + 0.5 * F[d]) / rho;
    double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    double s0 = L.w[0] * (-uu / (2.0 * L.cs2));
    double ugr = u[0] * grho[0] + u[1] * grho[1] + u[2] * grho[2];
    *p_o = L.cs2 / (1.0 - L.w[0]) * (s1 + 0.5 * ugr + rho * s0);
    *rho_o = rho;
}

int orc_mp_step(const orc_mp_params* m, double* f, double* h, double* phiu, int nsteps) {
    long N = (long)m->nx * m->ny * m->nz;
    lat_t L = make_lat(19);
    orc_params g;
    memset(&g, 0, sizeof(g));
    g.nx = m->nx; g.ny = m->ny; g.nz = m->nz;
    g.rho0 = 1.0;
    for (int k = 0; k < 6; k++) g.bc_f[k] = g.bc_g[k] = m->bc[k];
    orc_state st;
    memset(&st, 0, sizeof(st));
    int8_t* fl = calloc((size_t)N, 1);
    st.flag = fl;
    double* phi = malloc((size_t)N * sizeof(double));
    double* mu = malloc((size_t)N * sizeof(double));
    double* pf = malloc((size_t)19 * N * sizeof(double));
    double* ph_ = malloc((size_t)19 * N * sizeof(double));
    double* fresh = malloc((size_t)19 * N * sizeof(double));
    double eta = m->mobility / (L.cs2 * (m->tau_h - 0.5));
    for (int step = 0; step < nsteps; step++) {
        mp_phi_mu(m, h, phi, mu);
#pragma omp parallel for collapse(2) schedule(static)
        for (int X = 0; X < m->nx; X++)
            for (int Y = 0; Y < m->ny; Y++)
                for (int Z = 0; Z < m->nz; Z++) {
                    long id = lin(&g, X, Y, Z);
                    double fc[19], gphi[3], u[3], F[3], rho, p;
                    for (int i = 0; i < 19; i++) fc[i] = f[i * N + id];
                    mp_grad_lap(m, phi, X, Y, Z, gphi, NULL);
                    double ph = phi[id];
                    mp_macro_cell(m, fc, ph, mu[id], gphi, &rho, u, &p, F);
                    double grho[3];
                    for (int d = 0; d < 3; d++) grho[d] = (m->rho_l - m->rho_g) * gphi[d];
                    double pc = ph < 0 ? 0 : (ph > 1 ? 1 : ph);
                    double nu = mp_nu(m, pc);
                    double tf = nu / L.cs2 + 0.5;
                    double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
                    for (int i = 0; i < 19; i++) {
                        double cu = E27[i][0] * u[0] + E27[i][1] * u[1] + E27[i][2] * u[2];
                        double cg = E27[i][0] * grho[0] + E27[i][1] * grho[1] + E27[i][2] * grho[2];
                        double cF = E27[i][0] * F[0] + E27[i][1] * F[1] + E27[i][2] * F[2];
                        double si = L.w[i] * (cu / L.cs2 + cu * cu / (2 * L.cs2 * L.cs2) - uu / (2 * L.cs2));
                        double feq = (i == 0 ? (L.w[0] - 1.0) : L.w[i]) * p / L.cs2 + rho * si;
                        double Fi = (1.0 - 0.5 / tf) * L.w[i] * (cF + cu * cg) / L.cs2;
                        pf[i * N + id] = fc[i] - (fc[i] - feq) / tf + Fi;
                    }
                    /* Cahn-Hilliard lattice */
                    double nphiu[3] = {ph * u[0], ph * u[1], ph * u[2]};
                    double dpu[3];
                    for (int d = 0; d < 3; d++) dpu[d] = nphiu[d] - phiu[3 * id + d];
                    for (int i = 0; i < 19; i++) {
                        double cu = E27[i][0] * u[0] + E27[i][1] * u[1] + E27[i][2] * u[2];
                        double cd = E27[i][0] * dpu[0] + E27[i][1] * dpu[1] + E27[i][2] * dpu[2];
                        double heq = (i == 0) ? ph + (L.w[0] - 1.0) * eta * mu[id]
                                              : L.w[i] * eta * mu[id] + L.w[i] * ph * cu / L.cs2;
                        double Hi = (1.0 - 0.5 / m->tau_h) * L.w[i] * cd / L.cs2;
                        double hi = h[i * N + id];
                        ph_[i * N + id] = hi - (hi - heq) / m->tau_h + Hi;
                    }
                    for (int d = 0; d < 3; d++) phiu[3 * id + d] = nphiu[d];
                }
        stream(&g, &st, 19, m->bc, 0, NULL, NULL, NULL, NULL, pf, f, fresh);
        memcpy(f, fresh, (size_t)19 * N * sizeof(double));
        stream(&g, &st, 19, m->bc, 1, NULL, NULL, NULL, NULL, ph_, h, fresh);
        memcpy(h, fresh, (size_t)19 * N * sizeof(double));
    }
    free(fl); free(phi); free(mu); free(pf); free(ph_); free(fresh);
    return 0;
}

void orc_mp_macro(const orc_mp_params* m, const double* f, const double* h, double* phi,
                  double* mu, double* p, double* u3) {
    long N = (long)m->nx * m->ny * m->nz;
    orc_params g = {0};
    g.nx = m->nx; g.ny = m->ny; g.nz = m->nz;
    mp_phi_mu(m, h, phi, mu);
    for (int X = 0; X < m->nx; X++)
        for (int Y = 0; Y < m->ny; Y++)
            for (int Z = 0; Z < m->nz; Z++) {
                long id = lin(&g, X, Y, Z);
                double fc[19], gphi[3], F[3], rho;
                for (int i = 0; i < 19; i++) fc[i] = f[i * N + id];
                mp_grad_lap(m, phi, X, Y, Z, gphi, NULL);
                mp_macro_cell(m, fc, phi[id], mu[id], gphi, &rho, u3 + 3 * id, p + id, F);
            }
}

/* Equilibrium start from phi (N), u (3N or NULL), p (N or NULL); phiu receives phi u. */
void orc_mp_init(const orc_mp_params* m, const double* phi0, const double* u0, const double* p0,
                 double* f, double* h, double* phiu) {
    long N = (long)m->nx * m->ny * m->nz;
    lat_t L = make_lat(19);
    double* phi = malloc((size_t)N * sizeof(double));
    double* mu = malloc((size_t)N * sizeof(double));
    /* mu of the initial phi: h at rest equilibrium first, then mu from it */
    for (long x = 0; x < N; x++)
        for (int i = 0; i < 19; i++) h[i * N + x] = L.w[i] * phi0[x];
    mp_phi_mu(m, h, phi, mu);
    double eta = m->mobility / (L.cs2 * (m->tau_h - 0.5));
    for (long x = 0; x < N; x++) {
        double u[3] = {0, 0, 0};
        if (u0) for (int d = 0; d < 3; d++) u[d] = u0[3 * x + d];
        double ph = phi0[x], p = p0 ? p0[x] : 0.0;
        double rho = m->rho_g + ph * (m->rho_l - m->rho_g);
        double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        for (int i = 0; i < 19; i++) {
            double cu = E27[i][0] * u[0] + E27[i][1] * u[1] + E27[i][2] * u[2];
            double si = L.w[i] * (cu / L.cs2 + cu * cu / (2 * L.cs2 * L.cs2) - uu / (2 * L.cs2));
            f[i * N + x] = (i == 0 ? (L.w[0] - 1.0) : L.w[i]) * p / L.cs2 + rho * si;
            h[i * N + x] = (i == 0) ? ph + (L.w[0] - 1.0) * eta * mu[x]
                                    : L.w[i] * eta * mu[x] + L.w[i] * ph * cu / L.cs2;
        }
        for (int d = 0; d < 3; d++) phiu[3 * x + d] = ph * u[d];
    }
    free(phi);
    free(mu);
}

/* ------------------------------------------------------------------ flow adjoint */
/* Discrete adjoint of the flow-only step with the generalised moment collision
 * (ORC_OP_MRT form: momentum, traceless second order at s_nu, trace at s_bulk, odd and even
 * ghosts at s_odd / s_even; BGK and TRT are its special cases) and Guo forcing, for
 * L = 1/2 sum over non-wall cells |u_N - u*|^2 with u = (sum c f + F/2)/rho after nsteps.
 * Faces: periodic, bounce (any wall velocity: a constant term) or outflow (zero gradient);
 * wall cells.  grad[7] = dL/dF (3), dL/d(s_nu, s_bulk, s_odd, s_even).
 *
 * Reverse of one step, per non-wall cell y:
 *  - streaming transposed: post_i(y) was read by the cell y + c_i in direction i (no face
 *    crossed, y + c_i not a wall), or by y itself in direction inv(i) (bounce face or wall
 *    at y + c_i), and also by y in direction i when y - c_i lies behind an outflow face;
 *  - collision f* = feq + A (f - feq) + C G, A = sum_k (1 - r_k) P_k, C = sum_k (1 - r_k/2)
 *    P_k over the five parts; the projectors are w-orthogonal, P^T = W^-1 P W;
 *    lambda_f = A^T l + J_feq^T (I - A)^T l + J_G^T C^T l through rho = sum f and
 *    u = (j + F/2)/rho; dL/dr_k = -l . P_k (f - feq + G/2). */
int orc_flow_adjoint(const orc_params* p, const int8_t* flag, const double* f0, int nsteps,
                     const double* u_target, double* loss, double* grad) {
    int q = p->flow_q;
    if ((q != 19 && q != 27) || p->phase_q || p->thermal || p->solute || p->flow_incomp)
        return -1;
    for (int k = 0; k < 6; k++)
        if (p->bc_f[k] != 0 && p->bc_f[k] != 1 && p->bc_f[k] != 3) return -1;
    long N = NCELL(p);
    for (long x = 0; x < N; x++)
        if (flag[x] != 0 && flag[x] != 3) return -1; /* fluid and wall cells only */
    lat_t L = make_lat(q);
    mrt_t* M = (q == 19) ? &MRT19 : &MRT27;
    build_mrt(q, M);
    /* forward with a tape of the pre-collision populations */
    double* tape = malloc((size_t)(nsteps + 1) * q * N * sizeof(double));
    orc_state s;
    memset(&s, 0, sizeof(s));
    s.f = malloc((size_t)q * N * sizeof(double));
    s.flag = malloc((size_t)N);
    memcpy(s.flag, flag, (size_t)N);
    memcpy(s.f, f0, (size_t)q * N * sizeof(double));
    for (int n = 0; n < nsteps; n++) {
        memcpy(tape + (size_t)n * q * N, s.f, (size_t)q * N * sizeof(double));
        orc_step(p, &s, 1, ORC_OP_MRT, 1);
    }
    memcpy(tape + (size_t)nsteps * q * N, s.f, (size_t)q * N * sizeof(double));
    free(s.f);
    free(s.flag);
    for (int k = 0; k < 7; k++) grad[k] = 0;
    const double* F = p->force;
    /* loss and its adjoint */
    double* lam = calloc((size_t)q * N, sizeof(double));
    double* lst = calloc((size_t)q * N, sizeof(double));
    double Ls = 0;
    const double* fN = tape + (size_t)nsteps * q * N;
    for (long x = 0; x < N; x++) {
        if (flag[x] == 3) continue;
        double rho = 0, j[3] = {0, 0, 0};
        for (int i = 0; i < q; i++) {
            double v = fN[i * N + x];
            rho += v;
            for (int d = 0; d < 3; d++) j[d] += E27[i][d] * v;
        }
        double u[3], du[3];
        for (int d = 0; d < 3; d++) {
            u[d] = (j[d] + 0.5 * F[d]) / rho;
            du[d] = u[d] - u_target[3 * x + d];
            Ls += 0.5 * du[d] * du[d];
            grad[d] += du[d] / (2.0 * rho);
        }
        for (int i = 0; i < q; i++) {
            double s2 = 0;
            for (int d = 0; d < 3; d++) s2 += du[d] * (E27[i][d] - u[d]) / rho;
            lam[i * N + x] = s2;
        }
    }
    *loss = Ls;
    const double* Ps[5] = {M->mom, M->shear, M->bulk, M->godd, M->geven};
    const double rates[5] = {1.0, p->s_nu, p->s_bulk, p->s_odd, p->s_even};
    for (int n = nsteps - 1; n >= 0; n--) {
        const double* fn = tape + (size_t)n * q * N;
        /* streaming transposed: lst = adjoint of the post-collision populations */
        for (int x = 0; x < p->nx; x++)
            for (int y = 0; y < p->ny; y++)
                for (int z = 0; z < p->nz; z++) {
                    long id = lin(p, x, y, z);
                    if (flag[id] == 3) continue;
                    for (int i = 0; i < q; i++) {
                        int X = x + E27[i][0], Y = y + E27[i][1], Z = z + E27[i][2];
                        int face = resolve(p, p->bc_f, &X, &Y, &Z);
                        int k = inv_dir(i);
                        double v = 0;
                        if (face < 0) {
                            long t = lin(p, X, Y, Z);
                            v = flag[t] == 3 ? lam[k * N + id] : lam[i * N + t];
                        } else if (p->bc_f[face] == 1) {
                            v = lam[k * N + id];
                        }
                        int Xb = x - E27[i][0], Yb = y - E27[i][1], Zb = z - E27[i][2];
                        int fb = resolve(p, p->bc_f, &Xb, &Yb, &Zb);
                        if (fb >= 0 && p->bc_f[fb] == 3) v += lam[i * N + id];
                        lst[i * N + id] = v;
                    }
                }
        /* collision transposed */
        for (long x = 0; x < N; x++) {
            if (flag[x] == 3) {
                for (int i = 0; i < q; i++) lam[i * N + x] = 0;
                continue;
            }
            double f[27], l[27], lw[27], feq[27], G[27], ne[27], t[27];
            double rho = 0, j[3] = {0, 0, 0};
            for (int i = 0; i < q; i++) {
                f[i] = fn[i * N + x];
                l[i] = lst[i * N + x];
                lw[i] = L.w[i] * l[i];
                rho += f[i];
                for (int d = 0; d < 3; d++) j[d] += E27[i][d] * f[i];
            }
            double u[3];
            for (int d = 0; d < 3; d++) u[d] = (j[d] + 0.5 * F[d]) / rho;
            orc_equilibrium(q, rho, u, feq);
            guo_term(q, u, F, G);
            for (int i = 0; i < q; i++) ne[i] = f[i] - feq[i];
            double ATl[27] = {0}, CTl[27] = {0};
            for (int k = 0; k < 5; k++) {
                /* P^T l = W^-1 P (W l) */
                matvec(q, Ps[k], lw, t);
                for (int i = 0; i < q; i++) {
                    ATl[i] += (1.0 - rates[k]) * t[i] / L.w[i];
                    CTl[i] += (1.0 - 0.5 * rates[k]) * t[i] / L.w[i];
                }
                if (k > 0) {
                    double pn[27], pg[27], s1 = 0;
                    matvec(q, Ps[k], ne, pn);
                    matvec(q, Ps[k], G, pg);
                    for (int i = 0; i < q; i++) s1 += l[i] * (pn[i] + 0.5 * pg[i]);
                    grad[2 + k] -= s1;
                }
            }
            double rb = 0, ub[3] = {0, 0, 0};
            for (int i = 0; i < q; i++) {
                double mu_i = l[i] - ATl[i];  /* (I - A)^T l */
                double cu = E27[i][0] * u[0] + E27[i][1] * u[1] + E27[i][2] * u[2];
                double cF = E27[i][0] * F[0] + E27[i][1] * F[1] + E27[i][2] * F[2];
                rb += mu_i * feq[i] / rho;
                for (int d = 0; d < 3; d++) {
                    ub[d] += mu_i * L.w[i] * rho * (3.0 * E27[i][d] + 9.0 * cu * E27[i][d] - 3.0 * u[d]);
                    ub[d] += CTl[i] * L.w[i] * (-3.0 * F[d] + 9.0 * E27[i][d] * cF);
                    grad[d] += CTl[i] * L.w[i] * (3.0 * (E27[i][d] - u[d]) + 9.0 * cu * E27[i][d]);
                }
            }
            for (int d = 0; d < 3; d++) grad[d] += ub[d] / (2.0 * rho);
            for (int i = 0; i < q; i++) {
                double s2 = 0;
                for (int d = 0; d < 3; d++) s2 += ub[d] * (E27[i][d] - u[d]) / rho;
                lam[i * N + x] = ATl[i] + rb + s2;
            }
        }
    }
    free(lam);
    free(lst);
    free(tape);
    return 0;
}
