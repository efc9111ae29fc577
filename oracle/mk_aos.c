/* Reference-structured FP64 CPU step: the timing baseline for the fused step.
 *
 * TEST INFRASTRUCTURE ONLY (the CPU baseline of the benchmark and its parity test).
 *
 * The arithmetic is the oracle's (mk_oracle.c: the same cell functions and formulas, in the
 * same order, so the results agree bit for bit).  What differs is the data layout and the
 * sweep structure, which follow the reference's CPU path:
 *   - one array of node records laid out like mkNode (MaterKit/inc/mkNode.h:134-198), 3640 B
 *     per node, holding the _out / _hlp population pairs of every field
 *     (MaterKit/inc/mkFluidsLBM.h:95);
 *   - every operator is a separate OpenMP sweep over all nodes with schedule(static, 16):
 *     phase moments, flow relaxation, phase relaxation, thermal relaxation, and one pull
 *     propagation per field, out[i](x - e_i) -> hlp[i](x) (MaterKit/src/mkFluidsUser.cpp:
 *     68-134), with the neighbour found by division of the linear index (:73-75);
 *   - the per-step call order of run_flow_thermal_dendrite (MaterKit/MaterKit.cpp:500-518)
 *     with every field read at time level n, as in the fused step.
 * Scope: periodic faces, no wall cells, flow D3Q19 (or none), phase D3Q19 / D3Q7 (or none),
 * thermal D3Q7 (or none), no solute, no per-cell conductivity -- the benchmark workload.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "mk_oracle.h"

#define AOS_NODE_BYTES 3640

typedef struct aos_node {
    double f_out[19], f_hlp[19], f_equ[19];
    double g_out[19], g_hlp[19], g_nbr[19];
    double h_out[19], h_hlp[19];
    double x_out[19], x_hlp[19];
    double phi, phi_old, dphi, T, ns, rho;
    double vel[3];
    double X[4];
    int32_t flag, grain_id;
    unsigned char pad[AOS_NODE_BYTES - 10 * 19 * 8 - 6 * 8 - 3 * 8 - 4 * 8 - 2 * 4];
} aos_node;

_Static_assert(sizeof(aos_node) == AOS_NODE_BYTES, "node record size");

typedef struct {
    int q;
    int e[27 * 3];
    double w[27];
    int inv[27];
    double cs2;
} aos_lat;

static void aos_lattice(int q, aos_lat* L) {
    L->q = q;
    orc_lattice(q, L->e, L->w, L->inv, &L->cs2);
}

/* neighbour x - e of node id on the periodic box, by division of the linear index */
static long aos_upwind(const orc_params* p, long id, const int* e) {
    long nz = p->nz, ny = p->ny, nx = p->nx;
    long z = id % nz, y = (id / nz) % ny, x = id / (nz * ny);
    long X = (x - e[0] + nx) % nx, Y = (y - e[1] + ny) % ny, Z = (z - e[2] + nz) % nz;
    return (X * ny + Y) * nz + Z;
}

static long aos_at(const orc_params* p, long id, int dx, int dy, int dz) {
    int e[3] = {-dx, -dy, -dz};
    return aos_upwind(p, id, e);
}

static int aos_classify(const orc_params* p, int old, double phi) {
    if (old == 3 || old < 0 || p->pf_model != 0) return old;
    if (phi >= p->phi_s) return 2;
    if (phi > -p->phi_s) return 1;
    return 0;
}

static int aos_supported(const orc_params* p, const orc_state* s) {
    const int32_t* bcs[3] = {p->bc_f, p->bc_g, p->bc_h};
    for (int L = 0; L < 3; L++)
        for (int f = 0; f < 6; f++)
            if (bcs[L][f] != 0) return 0;
    if (p->flow_q != 0 && p->flow_q != 19) return 0;
    if (p->phase_q != 0 && p->phase_q != 7 && p->phase_q != 19) return 0;
    if (p->solute || (p->has_kappa && s->kappa)) return 0;
    long N = (long)p->nx * p->ny * p->nz;
    for (long x = 0; x < N; x++)
        if (s->flag[x] == 3) return 0;
    return 1;
}

static void aos_one_step(const orc_params* p, aos_node* nd, long N, int op, const aos_lat* Lf,
                         const aos_lat* Lg, const aos_lat* Lh) {
    /* phase moments: phi^n = sum g, flags */
    if (p->phase_q) {
#pragma omp parallel for schedule(static, 16)
        for (long id = 0; id < N; id++) {
            aos_node* n = &nd[id];
            n->phi_old = n->phi;
            double ph = 0;
            for (int i = 0; i < Lg->q; i++) ph += n->g_hlp[i];
            n->phi = ph;
            n->dphi = n->phi - n->phi_old;
            n->flag = aos_classify(p, n->flag, ph);
        }
    }
    /* macroscopic fields and flow relaxation */
#pragma omp parallel for schedule(static, 16)
    for (long id = 0; id < N; id++) {
        aos_node* n = &nd[id];
        if (p->thermal) {
            double T = 0;
            for (int i = 0; i < 7; i++) T += n->h_hlp[i];
            n->T = T;
        }
        double ns = 0;
        if (p->phase_q) {
            if (n->flag == 2) {
                ns = 1.0;
            } else if (n->flag == 1) {
                double fs = 0.5 * (1.0 + n->phi);
                ns = fs < 0 ? 0 : (fs > 1 ? 1 : fs);
            }
        }
        n->ns = ns;
        n->vel[0] = n->vel[1] = n->vel[2] = 0;
        if (p->flow_q) {
            double rho = 0, j[3] = {0, 0, 0}, F[3];
            for (int i = 0; i < Lf->q; i++) {
                double fi = n->f_hlp[i];
                rho += fi;
                for (int a = 0; a < 3; a++) j[a] += Lf->e[3 * i + a] * fi;
            }
            for (int a = 0; a < 3; a++) {
                F[a] = p->force[a];
                if (p->thermal) F[a] += p->rho0 * p->buoy[a] * (n->T - p->T_ref);
            }
            n->rho = rho;
            const double den = p->flow_incomp ? p->rho0 : rho;
            for (int a = 0; a < 3; a++) n->vel[a] = (j[a] + 0.5 * F[a]) / den;
            orc_collide_cell(p, op, n->f_hlp, F, n->f_out, ns);
        }
    }
    /* phase-field relaxation (gradient stencil from the neighbours' nodes) */
    if (p->phase_q) {
        double omg = 1.0 / (0.5 + p->pf_dt * p->pf_wdx2 / Lg->cs2);
#pragma omp parallel for schedule(static, 16)
        for (long id = 0; id < N; id++) {
            aos_node* n = &nd[id];
            double phi27[27];
            for (int dx = -1; dx <= 1; dx++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dz = -1; dz <= 1; dz++)
                        phi27[(dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)] =
                            nd[aos_at(p, id, dx, dy, dz)].phi;
            double upf = (p->pf_coupling == 1 && p->thermal) ? n->T : 0.0;
            double S, a;
            orc_phase_source(p, phi27, upf, n->grain_id, &S, &a);
            double ph = n->phi;
            double u[3];
            for (int k = 0; k < 3; k++) u[k] = (1.0 - n->ns) * n->vel[k];
            for (int i = 0; i < Lg->q; i++) {
                double cu = Lg->e[3 * i] * u[0] + Lg->e[3 * i + 1] * u[1] + Lg->e[3 * i + 2] * u[2];
                double geq = Lg->w[i] * (ph + (ph + 1.0) * cu / Lg->cs2);
                double gi = n->g_hlp[i];
                n->g_out[i] = gi - omg * (gi - geq) + Lg->w[i] * S;
            }
        }
    }
    /* thermal relaxation */
    if (p->thermal) {
#pragma omp parallel for schedule(static, 16)
        for (long id = 0; id < N; id++) {
            aos_node* n = &nd[id];
            double phi = p->phase_q ? n->phi : -1.0;
            double fs = 0.0;
            if (p->phase_q) {
                fs = 0.5 * (1.0 + phi);
                fs = fs < 0 ? 0 : (fs > 1 ? 1 : fs);
            }
            double alpha = p->alpha_l * (1.0 - fs) + p->alpha_s * fs;
            double om = 1.0 / (0.5 + alpha / Lh->cs2);
            double dphi = p->phase_q ? n->dphi : 0.0;
            double u[3];
            for (int k = 0; k < 3; k++) u[k] = (1.0 - n->ns) * n->vel[k];
            for (int i = 0; i < 7; i++) {
                double cu = Lh->e[3 * i] * u[0] + Lh->e[3 * i + 1] * u[1] + Lh->e[3 * i + 2] * u[2];
                double heq = Lh->w[i] * n->T * (1.0 + cu / Lh->cs2);
                double hi = n->h_hlp[i];
                n->h_out[i] = hi - om * (hi - heq) + Lh->w[i] * p->latent_T * dphi;
            }
        }
    }
    /* propagation: one pull sweep per field */
    if (p->flow_q) {
#pragma omp parallel for schedule(static, 16)
        for (long id = 0; id < N; id++)
            for (int i = 0; i < Lf->q; i++)
                nd[id].f_hlp[i] = nd[aos_upwind(p, id, &Lf->e[3 * i])].f_out[i];
    }
    if (p->phase_q) {
#pragma omp parallel for schedule(static, 16)
        for (long id = 0; id < N; id++)
            for (int i = 0; i < Lg->q; i++)
                nd[id].g_hlp[i] = nd[aos_upwind(p, id, &Lg->e[3 * i])].g_out[i];
    }
    if (p->thermal) {
#pragma omp parallel for schedule(static, 16)
        for (long id = 0; id < N; id++)
            for (int i = 0; i < 7; i++)
                nd[id].h_hlp[i] = nd[aos_upwind(p, id, &Lh->e[3 * i])].h_out[i];
    }
}

int orc_aos_step(const orc_params* p, orc_state* s, int op, int nsteps, double* seconds) {
    if (!aos_supported(p, s)) return -1;
    long N = (long)p->nx * p->ny * p->nz;
    aos_node* nd = (aos_node*)calloc((size_t)N, sizeof(aos_node));
    if (!nd) return -2;
    aos_lat Lf, Lg, Lh;
    aos_lattice(p->flow_q ? p->flow_q : 19, &Lf);
    aos_lattice(p->phase_q ? p->phase_q : 19, &Lg);
    aos_lattice(7, &Lh);
#pragma omp parallel for schedule(static, 16)
    for (long id = 0; id < N; id++) {
        aos_node* n = &nd[id];
        for (int i = 0; i < p->flow_q; i++) n->f_hlp[i] = s->f[i * N + id];
        for (int i = 0; i < p->phase_q; i++) n->g_hlp[i] = s->g[i * N + id];
        if (p->thermal)
            for (int i = 0; i < 7; i++) n->h_hlp[i] = s->h[i * N + id];
        if (p->phase_q) n->phi = s->phi[id];
        n->flag = s->flag[id];
        n->grain_id = s->grain ? s->grain[id] : 0;
    }
    double t0 = omp_get_wtime();
    for (int k = 0; k < nsteps; k++) aos_one_step(p, nd, N, op, &Lf, &Lg, &Lh);
    if (seconds) *seconds = omp_get_wtime() - t0;
#pragma omp parallel for schedule(static, 16)
    for (long id = 0; id < N; id++) {
        const aos_node* n = &nd[id];
        for (int i = 0; i < p->flow_q; i++) s->f[i * N + id] = n->f_hlp[i];
        for (int i = 0; i < p->phase_q; i++) s->g[i * N + id] = n->g_hlp[i];
        if (p->thermal)
            for (int i = 0; i < 7; i++) s->h[i * N + id] = n->h_hlp[i];
        if (p->phase_q) {
            s->phi[id] = n->phi;
            s->phi_prev[id] = n->phi_old;
        }
        s->flag[id] = (int8_t)n->flag;
    }
    free(nd);
    return 0;
}
