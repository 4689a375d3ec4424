/*
 * sph_oracle.c - CPU restatement of the reference SPH hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker for libsphgpu: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * It is never part of the product path.
 *
 * Each routine restates one reference routine, in the same order of
 * floating-point operations (so that, built without FMA contraction and
 * without reassociation, it reproduces the reference's bits):
 *   or_kernel          src/Numeric/kernel.f90:8-154
 *   or_bggs            src/NNPS/asymmetricSearch.f90:27-257
 *   kgc                src/Numeric/kernelGradientCorrection.f90:8-51, MCPE src/tools.f90:287-354
 *   sum_density        src/Physic/density.f90:10-80
 *   con_density        src/Physic/density.f90:83-107
 *   divergence         src/Numeric/divergence.f90:6-29
 *   internal_force     src/Physic/internalForce.f90:10-342 (SPH=2), eos.f90, viscosity.f90
 *   external_force     src/Physic/externalForce.f90:7-174
 *   arti_visc          src/Numeric/artificialViscosity.f90:7-114
 *   artificial_stress  src/Numeric/artificialTerm.f90:7-98
 *   h_upgrade          src/Numeric/hsml.f90:8-38
 *   single_step        src/TimeIntegration/singleStep.f90:47-144
 *   leap_frog          src/TimeIntegration/timeIntegration.f90:137-252
 *
 * Parity pinning (what checks each routine; see tests/test_oracle_*.py):
 *   or_bggs, or_kernel, con_density, internal_force (fluid + viscosity),
 *   external_force (gravity, wall LJ), arti_visc, leap_frog: bit/1e-13
 *   agreement with the compiled reference (oracle/_ref) on the reference
 *   dam break, and BGGS on random sets including window-miss cases.
 *   internal_force (solid), kgc, artificial_stress: compiled reference on the
 *   cantilever beams (can_beam, can_beam_3d).
 *   sum_density (+Shepard): compiled reference dam break with summation.
 *   h_upgrade (sle 1, 2): compiled reference dam break with sle switched on.
 *   The fluid_visc_skip_solid switch of internal_force has no reference
 *   counterpart (it drops a term the reference keeps); with it off, the
 *   routine is the reference's.
 *   external_force penalty and fluid-solid coupling branches, the wall
 *   force at the default range and the per-nick gravity variants: compiled
 *   reference on seeded custom sets (external_force alone, nicks unknown,
 *   db_gate, water_impact; tests/golden/custom_sets.py).
 *   leap_frog's J2 cap (type 104): compiled reference on seeded type-104
 *   blocks driven past the yield surface.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/sphgpu.h"

#define PI_D 3.141592653589793 /* acos(-1.0_8), tools.f90:8 */
#define GMR_M1 0.2000000476837158203125 /* real(1.2_4,8) - 1, asymmetricSearch.f90:21 */
#define NPG 3

/* ------------------------------------------------------------------ kernel */

/* kernel.f90:8-154. w and dwdx for one pair; dx has dim entries. */
static void or_kernel(int dim, int skf, double r, const double* dx, double hsml,
                      double* w, double* dwdx) {
    double q, factor = 0.0;
    int d;
    if (skf == 1) {
        q = r / hsml;
        if (dim == 1) factor = 1.0 / hsml;
        else if (dim == 2) factor = 15.0 / (7.0 * PI_D * (hsml * hsml));
        else if (dim == 3) factor = 3.0 / (2.0 * PI_D * (hsml * hsml * hsml));
        if (q >= 0 && q < 1) {
            *w = factor * (2.0 / 3.0 - q * q + (q * q * q) / 2.0);
            for (d = 0; d < dim; d++) dwdx[d] = factor * (-2.0 + 3.0 / 2.0 * q) / (hsml * hsml) * dx[d];
        } else if (q >= 1 && q <= 2) {
            double t = 2 - q;
            *w = factor * (1.0 / 6.0) * (t * t * t);
            for (d = 0; d < dim; d++) dwdx[d] = factor * (-1.0 / 2.0) * (t * t) / hsml * (dx[d] / r);
        } else {
            *w = 0;
            for (d = 0; d < dim; d++) dwdx[d] = 0;
        }
    } else if (skf == 2) {
        q = r / hsml;
        double sp = sqrt(PI_D) * hsml;
        factor = 1.0 / (dim == 1 ? sp : dim == 2 ? sp * sp : sp * sp * sp);
        if (q >= 0 && q < 3) {
            *w = factor * exp(-q * q);
            for (d = 0; d < dim; d++) dwdx[d] = *w * (-2.0 * dx[d] / (hsml * hsml));
        } else {
            *w = 0;
            for (d = 0; d < dim; d++) dwdx[d] = 0;
        }
    } else if (skf == 3) {
        q = r / hsml;
        if (dim == 1) factor = 1.0 / (120.0 * hsml);
        else if (dim == 2) factor = 7.0 / (478.0 * PI_D * (hsml * hsml));
        else if (dim == 3) factor = 1.0 / (120.0 * PI_D * (hsml * hsml * hsml));
        /* x**4 and x**5 are expanded left to right by the reference's
         * compiler: ((((x*x)*x)*x)*x) */
        double a = 3 - q, b = 2 - q, c = 1 - q;
        double a4 = a * a * a * a, b4 = b * b * b * b;
        double a5 = a4 * a, b5 = b4 * b, c5 = c * c * c * c * c;
        if (q >= 0 && q < 1) {
            *w = factor * (a5 - 6 * b5 + 15 * c5);
            for (d = 0; d < dim; d++) dwdx[d] = factor * (-120 + 120 * q - 50 * (q * q)) / (hsml * hsml) * dx[d];
        } else if (q >= 1 && q < 2) {
            *w = factor * (a5 - 6 * b5);
            for (d = 0; d < dim; d++) dwdx[d] = factor * (-5 * a4 + 30 * b4) / hsml * (dx[d] / r);
        } else if (q >= 2 && q <= 3) {
            *w = factor * a5;
            for (d = 0; d < dim; d++) dwdx[d] = factor * (-5 * a4) / hsml * (dx[d] / r);
        } else {
            *w = 0;
            for (d = 0; d < dim; d++) dwdx[d] = 0;
        }
    } else {
        *w = 0;
        for (d = 0; d < dim; d++) dwdx[d] = 0;
    }
}

int oracle_kernel(int dim, int skf, double r, const double* dx, double hsml, double* w,
                  double* dwdx) {
    or_kernel(dim, skf, r, dx, hsml, w, dwdx);
    return 0;
}

/* -------------------------------------------------------------------- BGGS */

typedef struct {
    int dim;
    double minc[3], maxc[3];
    int cellnum[3];
    int64_t ncell;
    int64_t* start; /* bucket start per cell (ncell+1) */
    int32_t* items; /* source ids, ascending within a bucket */
} or_grid;

/* asymmetricSearch.f90:27-47 (locate). Returns 0, or -1 if outside. */
static int or_locate(const or_grid* g, const double* x, int* cell) {
    for (int d = 0; d < g->dim; d++) {
        if (!(x[d] <= g->maxc[d]) || !(x[d] >= g->minc[d])) return -1;
        cell[d] = (int)(((x[d] - g->minc[d]) / (g->maxc[d] - g->minc[d])) * g->cellnum[d] + 1);
    }
    for (int d = g->dim; d < 3; d++) cell[d] = 1;
    return 0;
}

/* asymmetricSearch.f90:49-158 (initialize). */
static int or_grid_init(or_grid* g, int dim, int n, const double* x, const int* state) {
    g->dim = dim;
    for (int d = 0; d < 3; d++) {
        g->maxc[d] = 0; /* unused axes */
        g->minc[d] = 0;
        g->cellnum[d] = 1;
    }
    for (int d = 0; d < dim; d++) { g->maxc[d] = -1.7976931348623157e308; g->minc[d] = 1.7976931348623157e308; } /* -huge, huge */
    for (int i = 0; i < n; i++)
        for (int d = 0; d < dim; d++) {
            double xv = x[(int64_t)i * dim + d];
            if (xv > g->maxc[d]) g->maxc[d] = xv;
            if (xv < g->minc[d]) g->minc[d] = xv;
        }
    double delta[3] = {0, 0, 0};
    for (int d = 0; d < dim; d++) delta[d] = g->maxc[d] - g->minc[d];
    for (int d = 0; d < dim; d++) {
        g->maxc[d] = g->maxc[d] + GMR_M1 * delta[d];
        g->minc[d] = g->minc[d] - GMR_M1 * delta[d];
    }
    if (dim == 1) {
        int c = n / NPG + 1;
        g->cellnum[0] = c < 1000 ? c : 1000;
    } else if (dim == 2) {
        double t = pow(((double)n * delta[0]) / (delta[1] * NPG), 1.0 / 2);
        int c1 = (t >= 0 && t < 1e9) ? (int)t + 1 : 0;
        g->cellnum[0] = c1 < 1000 ? c1 : 1000;
        double t2 = g->cellnum[0] * delta[1] / delta[0];
        int c2 = (t2 >= 0 && t2 < 1e9) ? (int)t2 + 1 : 0;
        g->cellnum[1] = c2 < 1000 ? c2 : 1000;
    } else {
        double t = pow(((double)n * delta[0] * delta[0]) / (delta[1] * delta[2] * NPG), 1.0 / 3);
        int c1 = (t >= 0 && t < 1e9) ? (int)t + 1 : 0;
        g->cellnum[0] = c1 < 1000 ? c1 : 1000;
        double t2 = g->cellnum[0] * delta[1] / delta[0];
        double t3 = g->cellnum[0] * delta[2] / delta[0];
        int c2 = (t2 >= 0 && t2 < 1e9) ? (int)t2 + 1 : 0;
        int c3 = (t3 >= 0 && t3 < 1e9) ? (int)t3 + 1 : 0;
        g->cellnum[1] = c2 < 1000 ? c2 : 1000;
        g->cellnum[2] = c3 < 1000 ? c3 : 1000;
    }
    /* a flat bounding box (all sources on a line or plane) divides by zero
     * in the cell counts above; the reference's behaviour is undefined there */
    for (int d = 0; d < dim; d++)
        if (!(delta[d] > 0) || g->cellnum[d] < 1) {
            g->start = NULL;
            g->items = NULL;
            return SPHGPU_E_OUT_OF_RANGE;
        }
    g->ncell = (int64_t)g->cellnum[0] * g->cellnum[1] * g->cellnum[2];
    g->start = (int64_t*)calloc((size_t)g->ncell + 1, sizeof(int64_t));
    g->items = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t* cellof = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int bad = 0;
    for (int i = 0; i < n; i++) {
        cellof[i] = -1;
        if (state && state[i] == -1) continue;
        int c[3];
        if (or_locate(g, x + (int64_t)i * dim, c)) { bad = 1; continue; }
        int64_t lin = (int64_t)(c[0] - 1) + (int64_t)g->cellnum[0] * ((c[1] - 1) + (int64_t)g->cellnum[1] * (c[2] - 1));
        cellof[i] = (int32_t)lin;
        g->start[lin + 1]++;
    }
    for (int64_t c = 0; c < g->ncell; c++) g->start[c + 1] += g->start[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)g->ncell);
    memcpy(fill, g->start, sizeof(int64_t) * (size_t)g->ncell);
    for (int i = 0; i < n; i++)
        if (cellof[i] >= 0) g->items[fill[cellof[i]]++] = i;
    free(fill);
    free(cellof);
    return bad ? SPHGPU_E_OUT_OF_RANGE : 0;
}

static void or_grid_free(or_grid* g) {
    free(g->start);
    free(g->items);
}

typedef struct {
    int64_t* off;  /* ntotal+1 */
    int32_t* nbr;
    double* w;
    double* dwdx;  /* dim per pair */
    int64_t npairs;
} or_pairs;

static void or_pairs_free(or_pairs* p) {
    free(p->off); free(p->nbr); free(p->w); free(p->dwdx);
    memset(p, 0, sizeof(*p));
}

/* asymmetricSearch.f90:160-257 (BGGS). Targets 0..ntotal-1, sources 0..n-1,
 * skipItsSelf = true. Pairs in the reference's order (z, y, x cell loops,
 * bucket order). */
static int or_bggs(int dim, int skf, int ntotal, int n, const double* x, const double* h,
                   const int* state, or_pairs* out, or_grid* gout) {
    int scale_k = (skf == 1) ? 2 : (skf == 2 || skf == 3) ? 3 : 1;
    or_grid g;
    int rc = or_grid_init(&g, dim, n, x, state);
    if (rc) { or_grid_free(&g); return rc; }
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    /* per-thread growable buffers over contiguous target chunks */
    int32_t** tn = (int32_t**)calloc(nthr, sizeof(void*));
    double** tw = (double**)calloc(nthr, sizeof(void*));
    double** td = (double**)calloc(nthr, sizeof(void*));
    int64_t* tcount = (int64_t*)calloc(nthr, sizeof(int64_t));
    int64_t* cnt = (int64_t*)calloc((size_t)ntotal + 1, sizeof(int64_t));
    int err = 0;
#pragma omp parallel num_threads(nthr) reduction(| : err)
    {
        int t = 0, nt = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        nt = omp_get_num_threads();
#endif
        int64_t lo = (int64_t)ntotal * t / nt, hi = (int64_t)ntotal * (t + 1) / nt;
        int64_t cap = 1024, m = 0;
        int32_t* bn = (int32_t*)malloc(sizeof(int32_t) * cap);
        double* bw = (double*)malloc(sizeof(double) * cap);
        double* bd = (double*)malloc(sizeof(double) * cap * dim);
        for (int64_t i = lo; i < hi; i++) {
            if (state && state[i] == -1) continue;
            const double* xi = x + i * dim;
            int cell[3], wc[3], mn[3] = {1, 1, 1}, mx[3] = {1, 1, 1};
            if (or_locate(&g, xi, cell)) { err |= 1; continue; }
            for (int d = 0; d < dim; d++) {
                wc[d] = (int)(h[i] / (g.maxc[d] - g.minc[d]) * g.cellnum[d]) + 1;
                mx[d] = cell[d] + wc[d] < g.cellnum[d] ? cell[d] + wc[d] : g.cellnum[d];
                mn[d] = cell[d] - wc[d] > 1 ? cell[d] - wc[d] : 1;
            }
            int64_t c0 = m;
            for (int zc = mn[2]; zc <= mx[2]; zc++)
                for (int yc = mn[1]; yc <= mx[1]; yc++)
                    for (int xc = mn[0]; xc <= mx[0]; xc++) {
                        int64_t lin = (int64_t)(xc - 1) + (int64_t)g.cellnum[0] * ((yc - 1) + (int64_t)g.cellnum[1] * (zc - 1));
                        for (int64_t k = g.start[lin]; k < g.start[lin + 1]; k++) {
                            int j = g.items[k];
                            if (state && state[j] == -1) continue;
                            if (i == j) continue;
                            double dx[3];
                            dx[0] = xi[0] - x[(int64_t)j * dim];
                            double dr = dx[0] * dx[0];
                            for (int d = 1; d < dim; d++) {
                                dx[d] = xi[d] - x[(int64_t)j * dim + d];
                                dr = dr + dx[d] * dx[d];
                            }
                            double r = sqrt(dr);
                            double mh = h[i];
                            if (r < mh * scale_k) {
                                if (m == cap) {
                                    cap *= 2;
                                    bn = (int32_t*)realloc(bn, sizeof(int32_t) * cap);
                                    bw = (double*)realloc(bw, sizeof(double) * cap);
                                    bd = (double*)realloc(bd, sizeof(double) * cap * dim);
                                }
                                bn[m] = j;
                                or_kernel(dim, skf, r, dx, mh, &bw[m], &bd[m * dim]);
                                m++;
                            }
                        }
                    }
            cnt[i + 1] = m - c0;
        }
        tn[t] = bn; tw[t] = bw; td[t] = bd; tcount[t] = m;
    }
    out->off = (int64_t*)malloc(sizeof(int64_t) * ((size_t)ntotal + 1));
    out->off[0] = 0;
    for (int64_t i = 0; i < ntotal; i++) out->off[i + 1] = out->off[i] + cnt[i + 1];
    out->npairs = out->off[ntotal];
    out->nbr = (int32_t*)malloc(sizeof(int32_t) * (size_t)(out->npairs + 1));
    out->w = (double*)malloc(sizeof(double) * (size_t)(out->npairs + 1));
    out->dwdx = (double*)malloc(sizeof(double) * (size_t)(out->npairs + 1) * dim);
    int64_t pos = 0;
    for (int t = 0; t < nthr; t++) {
        if (tcount[t]) {
            memcpy(out->nbr + pos, tn[t], sizeof(int32_t) * tcount[t]);
            memcpy(out->w + pos, tw[t], sizeof(double) * tcount[t]);
            memcpy(out->dwdx + pos * dim, td[t], sizeof(double) * tcount[t] * dim);
        }
        pos += tcount[t];
        free(tn[t]); free(tw[t]); free(td[t]);
    }
    free(tn); free(tw); free(td); free(tcount); free(cnt);
    if (gout) *gout = g; else or_grid_free(&g);
    return err ? SPHGPU_E_OUT_OF_RANGE : 0;
}

/* Exported BGGS: fills CSR arrays (caller sizes nbr/w/dwdx with cap pairs).
 * grid9 (optional) gets minCoor, maxCoor, cellNum. */
int oracle_bggs(int dim, int skf, int ntotal, int n, const double* x, const double* h,
                const int32_t* state, int64_t cap, int64_t* offsets, int32_t* nbr, double* w,
                double* dwdx, int64_t* needed, double* grid9) {
    or_pairs p;
    or_grid g;
    int rc = or_bggs(dim, skf, ntotal, n, x, h, state, &p, &g);
    if (rc) return rc;
    if (grid9)
        for (int d = 0; d < 3; d++) {
            grid9[d] = g.minc[d]; grid9[3 + d] = g.maxc[d]; grid9[6 + d] = g.cellnum[d];
        }
    or_grid_free(&g);
    if (needed) *needed = p.npairs;
    if (p.npairs > cap) { or_pairs_free(&p); return SPHGPU_E_PAIR_CAPACITY; }
    memcpy(offsets, p.off, sizeof(int64_t) * ((size_t)ntotal + 1));
    if (nbr) memcpy(nbr, p.nbr, sizeof(int32_t) * p.npairs);
    if (w) memcpy(w, p.w, sizeof(double) * p.npairs);
    if (dwdx) memcpy(dwdx, p.dwdx, sizeof(double) * p.npairs * dim);
    or_pairs_free(&p);
    return 0;
}

/* ------------------------------------------------------------ state / step */

typedef struct {
    int dim, ntotal, n;
    double *x, *v, *disp, *mass, *rho, *p, *e, *c, *h, *eta, *divx, *divv, *stress;
    const int32_t *type, *state, *boundary;
    /* integrator state (timeIntegration.f90:41-46, :362-368) */
    double *shear, *shear_prev, *dsdt, *v_prev, *rho_prev, *e_prev;
    double *dvdt, *drhodt, *dedt;
} or_state;

static const sphgpu_material* or_mat(const sphgpu_config* cfg, int type) {
    int a = type < 0 ? -type : type;
    for (int k = 0; k < cfg->n_materials && k < SPHGPU_MAX_MATERIALS; k++)
        if (cfg->mat[k].type == a) return &cfg->mat[k];
    return NULL;
}

/* eos.f90:23-53 (Tait), :149-157 / :219-248 (linear), :57-64 (isothermal). */
static void or_eos(const sphgpu_material* m, double rho, double* p) {
    if (!m) return;
    switch (m->eos) {
    case SPHGPU_EOS_TAIT: *p = m->K * (pow(rho / m->rho0, m->gamma) - 1); break;
    case SPHGPU_EOS_LINEAR: *p = m->K * (rho - m->rho0); break;
    case SPHGPU_EOS_ISOTHERMAL: *p = m->K * rho; break;
    default: break;
    }
}

#define IDX2(i, d, e) (((int64_t)(i) * dim + (d)) * dim + (e))

/* MCPE, tools.f90:287-354. A is dim x dim, row-major A[r*n+c] = A(r,c). */
static void or_mcpe(int n, double* A, double* B, double* X) {
    double buffer, temp;
    for (int i = 0; i < n; i++) {
        int p = i;
        for (int j = i + 1; j < n; j++)
            if (fabs(A[p * n + i]) < fabs(A[j * n + i])) p = j;
        for (int j = 0; j < n; j++) {
            buffer = A[i * n + j]; A[i * n + j] = A[p * n + j]; A[p * n + j] = buffer;
        }
        buffer = B[i]; B[i] = B[p]; B[p] = buffer;
    }
    for (int i = 0; i < n - 1; i++)
        for (int j = i + 1; j < n; j++) {
            buffer = A[j * n + i] / A[i * n + i];
            temp = 0;
            for (int k = i; k < n; k++) {
                A[j * n + k] = A[j * n + k] - A[i * n + k] * buffer;
                temp = temp + A[j * n + k] * A[j * n + k];
            }
            B[j] = B[j] - B[i] * buffer;
            if (temp == 0) return;
        }
    X[n - 1] = B[n - 1] / A[(n - 1) * n + (n - 1)];
    for (int i = n - 2; i >= 0; i--) {
        for (int j = n - 1; j >= i + 1; j--) B[i] = B[i] - A[i * n + j] * X[j];
        X[i] = B[i] / A[i * n + i];
    }
}

/* kernelGradientCorrection.f90:8-51 */
static void or_kgc(const or_state* s, or_pairs* pr) {
    int dim = s->dim;
    double* A = (double*)calloc((size_t)s->ntotal * dim * dim, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->ntotal; i++) {
        if (s->type[i] <= 100) continue;
        for (int64_t k = pr->off[i]; k < pr->off[i + 1]; k++) {
            int j = pr->nbr[k];
            int ai = s->type[i] < 0 ? -s->type[i] : s->type[i];
            int aj = s->type[j] < 0 ? -s->type[j] : s->type[j];
            if ((ai - 100) * (aj - 100) < 0) continue;
            double dx[3];
            for (int d = 0; d < dim; d++) dx[d] = s->x[(int64_t)j * dim + d] - s->x[(int64_t)i * dim + d];
            for (int d = 0; d < dim; d++)
                for (int e = 0; e < dim; e++)
                    A[IDX2(i, d, e)] = A[IDX2(i, d, e)] + dx[e] * pr->dwdx[k * dim + d] * s->mass[j] / s->rho[j];
        }
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->ntotal; i++) {
        for (int64_t k = pr->off[i]; k < pr->off[i + 1]; k++) {
            double T[9], B[3];
            memcpy(T, A + (int64_t)i * dim * dim, sizeof(double) * dim * dim);
            memcpy(B, pr->dwdx + k * dim, sizeof(double) * dim);
            or_mcpe(dim, T, B, pr->dwdx + k * dim);
        }
    }
    free(A);
}

/* density.f90:10-80. Shepard weights use divergencePosition of the
 * previous evaluation, as the reference does. */
static void or_sum_density(const sphgpu_config* cfg, or_state* s, const or_pairs* pr) {
    int dim = s->dim, N = s->ntotal;
    double hv[3] = {0, 0, 0}, self, dummy[3];
    double* wi = (double*)calloc((size_t)N + 1, sizeof(double));
    if (cfg->normalize_density) {
        for (int i = 0; i < N; i++)
            if (s->divx[i] < 1.5) {
                or_kernel(dim, cfg->skf, 0.0, hv, s->h[i], &self, dummy);
                wi[i] = s->mass[i] / s->rho[i] * self;
            }
        /* REDUCTION(+:wi) (density.f90:47): the neighbour sum is formed in a
         * zero-initialised private copy and then added to the self term */
        for (int i = 0; i < N; i++)
            if (s->divx[i] < 1.5) {
                double acc = 0;
                for (int64_t k = pr->off[i]; k < pr->off[i + 1]; k++) {
                    int j = pr->nbr[k];
                    acc = acc + s->mass[j] / s->rho[j] * pr->w[k];
                }
                wi[i] = wi[i] + acc;
            }
    }
    for (int i = 0; i < N; i++) {
        or_kernel(dim, cfg->skf, 0.0, hv, s->h[i], &self, dummy);
        s->rho[i] = s->mass[i] * self;
    }
    for (int i = 0; i < N; i++)
        for (int64_t k = pr->off[i]; k < pr->off[i + 1]; k++) {
            int j = pr->nbr[k];
            s->rho[i] = s->rho[i] + s->mass[j] * pr->w[k];
        }
    if (cfg->normalize_density)
        for (int i = 0; i < N; i++)
            if (s->divx[i] < 1.5) s->rho[i] = s->rho[i] / wi[i];
    free(wi);
}

/* density.f90:83-107 */
static void or_con_density(or_state* s, const or_pairs* pr, double* drhodt) {
    int dim = s->dim;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->ntotal; i++) {
        double acc = 0;
        for (int64_t k = pr->off[i]; k < pr->off[i + 1]; k++) {
            int j = pr->nbr[k];
            double dot = 0;
            for (int d = 0; d < dim; d++)
                dot = dot + (s->v[(int64_t)i * dim + d] - s->v[(int64_t)j * dim + d]) * pr->dwdx[k * dim + d];
            acc = acc + s->mass[j] * dot * s->rho[i] / s->rho[j];
        }
        drhodt[i] = acc;
    }
}

/* divergence.f90:6-29 */
static void or_divergence(or_state* s, const or_pairs* pr) {
    int dim = s->dim;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->ntotal; i++) {
        double ax = 0, av = 0;
        for (int64_t k = pr->off[i]; k < pr->off[i + 1]; k++) {
            int j = pr->nbr[k];
            double sx = 0, sv = 0;
            for (int d = 0; d < dim; d++)
                sx = sx + (s->x[(int64_t)j * dim + d] - s->x[(int64_t)i * dim + d]) * pr->dwdx[k * dim + d];
            for (int d = 0; d < dim; d++)
                sv = sv + (s->v[(int64_t)j * dim + d] - s->v[(int64_t)i * dim + d]) * pr->dwdx[k * dim + d];
            ax = ax + s->mass[j] / s->rho[j] * sx;
            av = av + s->mass[j] / s->rho[j] * sv;
        }
        s->divx[i] = ax;
        s->divv[i] = av;
    }
}

/* Unit vector from i to j (norm2 of the separation) and the Riemann
 * interface state of density.f90:120-141 / internalForce.f90:285-305.
 * Z_l, Z_r: acoustic impedances; pi, pj: the pressures it is given. */
static void or_riemann(const or_state* s, int i, int j, double Z_l, double Z_r, double pi, double pj,
                       double* e, double* v_l, double* v_r, double* v_ij, double* v_star) {
    int dim = s->dim;
    const double* xi = s->x + (int64_t)i * dim;
    const double* xj = s->x + (int64_t)j * dim;
    const double* vi = s->v + (int64_t)i * dim;
    const double* vj = s->v + (int64_t)j * dim;
    double ss = 0;
    for (int d = 0; d < dim; d++) ss = ss + (xj[d] - xi[d]) * (xj[d] - xi[d]);
    double nrm = sqrt(ss);
    for (int d = 0; d < dim; d++) e[d] = (xj[d] - xi[d]) / nrm;
    double a = 0, b = 0;
    for (int d = 0; d < dim; d++) { a = a + vi[d] * e[d]; b = b + vj[d] * e[d]; }
    *v_l = a;
    *v_r = b;
    *v_ij = (Z_l * a + Z_r * b + (pi - pj)) / (Z_l + Z_r);
    for (int d = 0; d < dim; d++)
        v_star[d] = *v_ij * e[d] + ((vi[d] + vj[d]) / 2 - ((a + b) / 2) * e[d]);
}

/* Fluid pair sums of SPH = 1 (internalForce.f90:197-225) and SPH = 3, the
 * PVRS Riemann form (:277-325), for particle i. */
static void or_pair_sph13(const sphgpu_config* cfg, const or_state* s, const or_pairs* pr,
                          const double* edot, int i, double* acc, double* de) {
    int dim = s->dim;
    const double* vi = s->v + (int64_t)i * dim;
    const double* Ei = edot + (int64_t)i * dim * dim;
    for (int64_t k = pr->off[i]; k < pr->off[i + 1]; k++) {
        int j = pr->nbr[k];
        const double* dw = pr->dwdx + k * dim;
        const double* vj = s->v + (int64_t)j * dim;
        const double* Ej = edot + (int64_t)j * dim * dim;
        double rhoij = 1.0 / (s->rho[i] * s->rho[j]);
        double aux, dv[3];
        if (cfg->pa_sph == 1) {
            aux = (s->p[i] + s->p[j]) * rhoij;
            for (int d = 0; d < dim; d++) dv[d] = vi[d] - vj[d];
        } else {
            double Z_l = s->rho[i] * s->c[i], Z_r = s->rho[j] * s->c[j];
            double e[3], v_l, v_r, v_ij, v_star[3];
            or_riemann(s, i, j, Z_l, Z_r, s->p[i], s->p[j], e, &v_l, &v_r, &v_ij, v_star);
            double p_star = (Z_l * s->p[j] + Z_r * s->p[i] + Z_l * Z_r * (v_l - v_r)) / (Z_l + Z_r);
            aux = 2 * p_star * rhoij;
            for (int d = 0; d < dim; d++) dv[d] = 2 * (vi[d] - v_star[d]);
        }
        for (int d = 0; d < dim; d++) {
            double mv = 0;   /* matmul(eta_i edot_i + eta_j edot_j, dwdx) */
            for (int e = 0; e < dim; e++)
                mv = mv + (s->eta[i] * Ei[d * dim + e] + s->eta[j] * Ej[d * dim + e]) * dw[e];
            acc[d] = acc[d] + s->mass[j] * (-aux * dw[d] + mv * rhoij);
        }
        double dot = 0;
        for (int d = 0; d < dim; d++) dot = dot + dv[d] * dw[d];
        *de = *de + s->mass[j] * aux * dot;
    }
}

/* density.f90:109-145, PVRS Riemann solver. Quirk kept: Z_l takes the
 * neighbour's sound speed (P(j)%SoundSpeed), and the pressures are the ones
 * stored before this step's EOS. */
static void or_con_density_riemann(or_state* s, const or_pairs* pr, double* drhodt) {
    int dim = s->dim;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->ntotal; i++) {
        double acc = 0;
        const double* vi = s->v + (int64_t)i * dim;
        for (int64_t k = pr->off[i]; k < pr->off[i + 1]; k++) {
            int j = pr->nbr[k];
            double Z_l = s->rho[i] * s->c[j], Z_r = s->rho[j] * s->c[j];
            double e[3], v_l, v_r, v_ij, v_star[3];
            or_riemann(s, i, j, Z_l, Z_r, s->p[i], s->p[j], e, &v_l, &v_r, &v_ij, v_star);
            double dot = 0;
            for (int d = 0; d < dim; d++) dot = dot + (vi[d] - v_star[d]) * pr->dwdx[k * dim + d];
            acc = acc + 2 * s->rho[i] * s->mass[j] / s->rho[j] * dot;
        }
        drhodt[i] = acc;
    }
}

/* internalForce.f90:10-342 with SOLID=1 and Shear present, SPH = 1, 2, 3. */
static void or_internal_force(const sphgpu_config* cfg, or_state* s, const or_pairs* pr,
                              double* dvdt, double* dedt) {
    int dim = s->dim, N = s->n, nt = s->ntotal;
    double* tdsdt = (double*)calloc((size_t)N, sizeof(double));
    double* edot = (double*)calloc((size_t)N * dim * dim, sizeof(double));
    double* rdot = (double*)calloc((size_t)N * dim * dim, sizeof(double));
    if (cfg->viscosity)
        for (int i = 0; i < N; i++) {
            const sphgpu_material* m = or_mat(cfg, s->type[i]);
            s->eta[i] = m ? m->eta : 0.0; /* viscosity.f90:14-33 */
        }
    if (cfg->viscosity) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nt; i++) { /* particles >= ntotal have no neighbours */
            int ai = s->type[i] < 0 ? -s->type[i] : s->type[i];
            double* E = edot + (int64_t)i * dim * dim;
            double* R = rdot + (int64_t)i * dim * dim;
            for (int64_t k = pr->off[i]; k < pr->off[i + 1]; k++) {
                int j = pr->nbr[k];
                const double* dw = pr->dwdx + k * dim;
                double dv[3];
                for (int d = 0; d < dim; d++) dv[d] = s->v[(int64_t)j * dim + d] - s->v[(int64_t)i * dim + d];
                if (ai <= 100) {
                    for (int d = 0; d < dim; d++)
                        for (int dd = 0; dd < dim; dd++) {
                            E[d * dim + dd] = E[d * dim + dd] + s->mass[j] / s->rho[j] * dv[d] * dw[dd];
                            E[dd * dim + d] = E[dd * dim + d] + s->mass[j] / s->rho[j] * dv[dd] * dw[d];
                            if (d == dd)
                                for (int ddd = 0; ddd < dim; ddd++)
                                    E[d * dim + d] = E[d * dim + d] - 2.0 / 3 * s->mass[j] / s->rho[j] * dv[ddd] * dw[ddd];
                        }
                } else {
                    for (int d = 0; d < dim; d++)
                        for (int dd = 0; dd < dim; dd++) {
                            double vab = 0.5 * s->mass[j] / s->rho[j] * dv[d] * dw[dd];
                            double vba = 0.5 * s->mass[j] / s->rho[j] * dv[dd] * dw[d];
                            E[d * dim + dd] = E[d * dim + dd] + (vab + vba);
                            R[d * dim + dd] = R[d * dim + dd] + (vab - vba);
                        }
                }
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        int ai = s->type[i] < 0 ? -s->type[i] : s->type[i];
        const sphgpu_material* m = or_mat(cfg, s->type[i]);
        double* E = edot + (int64_t)i * dim * dim;
        if (ai <= 100) {
            if (cfg->viscosity) {
                double sq = 0; /* sum(edot(:,:,i)**2), column-major element order */
                for (int e = 0; e < dim; e++)
                    for (int d = 0; d < dim; d++) sq = sq + E[d * dim + e] * E[d * dim + e];
                tdsdt[i] = 0.5 * s->eta[i] / s->rho[i] * sq;
            }
            or_eos(m, s->rho[i], &s->p[i]);
        } else {
            const double* S = s->shear + (int64_t)i * dim * dim;
            const double* R = rdot + (int64_t)i * dim * dim;
            double aux = 0, ae[9], sq = 0;
            for (int d = 0; d < dim; d++) aux = aux + E[d * dim + d];
            aux = aux / 3;
            memcpy(ae, E, sizeof(double) * dim * dim);
            for (int d = 0; d < dim; d++) ae[d * dim + d] = ae[d * dim + d] - aux;
            for (int e = 0; e < dim; e++)
                for (int d = 0; d < dim; d++) sq = sq + S[d * dim + e] * ae[d * dim + e];
            tdsdt[i] = sq / s->rho[i];
            or_eos(m, s->rho[i], &s->p[i]);
            if (i < nt) {
                double* D = s->dsdt + (int64_t)i * dim * dim;
                for (int d = 0; d < dim; d++)
                    for (int dd = 0; dd < dim; dd++) {
                        D[d * dim + dd] = 2 * s->eta[i] * ae[d * dim + dd];
                        for (int ddd = 0; ddd < dim; ddd++)
                            D[d * dim + dd] = D[d * dim + dd] + S[d * dim + ddd] * R[dd * dim + ddd] + S[dd * dim + ddd] * R[d * dim + ddd];
                    }
            }
            double* St = s->stress + (int64_t)i * dim * dim;
            for (int d = 0; d < dim; d++) {
                for (int dd = 0; dd < dim; dd++) St[d * dim + dd] = S[d * dim + dd];
                St[d * dim + d] = St[d * dim + d] - s->p[i];
            }
        }
    }
    /* SPH = 2 pair pass, internalForce.f90:228-274 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nt; i++) {
        int ai = s->type[i] < 0 ? -s->type[i] : s->type[i];
        double acc[3] = {0, 0, 0}, de = 0;
        const double* vi = s->v + (int64_t)i * dim;
        double ri2 = s->rho[i] * s->rho[i];
        if (cfg->pa_sph != 2) {
            /* SPH = 1 (:197-225) and 3 (:277-325): fluids only */
            if (ai <= 100) or_pair_sph13(cfg, s, pr, edot, i, acc, &de);
        } else if (ai <= 100) {
            const double* Ei = edot + (int64_t)i * dim * dim;
            for (int64_t k = pr->off[i]; k < pr->off[i + 1]; k++) {
                int j = pr->nbr[k];
                const double* dw = pr->dwdx + k * dim;
                const double* vj = s->v + (int64_t)j * dim;
                const double* Ej = edot + (int64_t)j * dim * dim;
                double rj2 = s->rho[j] * s->rho[j];
                double aux = s->p[i] / ri2 + s->p[j] / rj2;
                const int skip = cfg->fluid_visc_skip_solid && abs(s->type[j]) > 100;
                for (int d = 0; d < dim; d++) {
                    double mv = 0;
                    for (int e = 0; e < dim && !skip; e++)
                        mv = mv + (s->eta[i] * Ei[d * dim + e] / ri2 + s->eta[j] * Ej[d * dim + e] / rj2) * dw[e];
                    acc[d] = acc[d] + s->mass[j] * (-aux * dw[d] + mv);
                }
                double dot = 0;
                for (int d = 0; d < dim; d++) dot = dot + (vi[d] - vj[d]) * dw[d];
                de = de + s->mass[j] * aux * dot;
            }
        } else {
            const double* Si = s->stress + (int64_t)i * dim * dim;
            for (int64_t k = pr->off[i]; k < pr->off[i + 1]; k++) {
                int j = pr->nbr[k];
                const double* dw = pr->dwdx + k * dim;
                const double* vj = s->v + (int64_t)j * dim;
                const double* Sj = s->stress + (int64_t)j * dim * dim;
                double rj2 = s->rho[j] * s->rho[j];
                for (int d = 0; d < dim; d++) {
                    for (int dd = 0; dd < dim; dd++)
                        acc[d] = acc[d] + s->mass[j] * (Si[d * dim + dd] / ri2 + Sj[d * dim + dd] / rj2) * dw[dd];
                    de = de + s->mass[j] * (s->p[i] / ri2 + s->p[j] / rj2) * (vi[d] - vj[d]) * dw[d];
                }
            }
        }
        for (int d = 0; d < dim; d++) dvdt[(int64_t)i * dim + d] = acc[d];
        dedt[i] = tdsdt[i] + 0.5 * de;
    }
    free(tdsdt); free(edot); free(rdot);
}

/* externalForce.f90:7-174 */
static void or_external_force(const sphgpu_config* cfg, or_state* s, const or_pairs* pr,
                              double* dvdt) {
    int dim = s->dim, nt = s->ntotal;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nt; i++) {
        double acc[3] = {0, 0, 0};
        if (cfg->gravity) {
            if (!(cfg->gravity_mode == SPHGPU_GRAVITY_SKIP_SOLID && s->type[i] > 100))
                acc[dim - 1] = cfg->gravity_g;
        }
        const double* xi = s->x + (int64_t)i * dim;
        int ti = s->type[i];
        for (int64_t k = pr->off[i]; k < pr->off[i + 1]; k++) {
            int j = pr->nbr[k];
            int tj = s->type[j];
            const double* xj = s->x + (int64_t)j * dim;
            double dx[3], dr, r;
            if (ti > 0 && tj < 0) {
                dx[0] = xi[0] - xj[0];
                dr = dx[0] * dx[0];
                for (int d = 1; d < dim; d++) { dx[d] = xi[d] - xj[d]; dr = dr + dx[d] * dx[d]; }
                r = sqrt(dr);
                if (r < cfg->lj_r0) {
                    double f = cfg->lj_factor * (pow(cfg->lj_r0 / r, cfg->lj_p1) - pow(cfg->lj_r0 / r, cfg->lj_p2));
                    for (int d = 0; d < dim; d++) acc[d] = acc[d] + f * dx[d] / (r * r);
                }
            }
            if (ti != tj && ti > 0 && tj > 0) {
                if (ti < 100 && tj < 100) {
                    dx[0] = xi[0] - xj[0];
                    dr = dx[0] * dx[0];
                    for (int d = 1; d < dim; d++) { dx[d] = xi[d] - xj[d]; dr = dr + dx[d] * dx[d]; }
                    r = sqrt(dr);
                    double pe = (s->h[i] + s->h[j]) / (2 * r);
                    if (pe >= 1) {
                        double f = cfg->pen_factor * (pow(pe, cfg->pen_n1) - pow(pe, cfg->pen_n2));
                        for (int d = 0; d < dim; d++) acc[d] = acc[d] + f * dx[d] / (r * r);
                    }
                } else if (!(ti > 100 && tj > 100)) {
                    dx[0] = xi[0] - xj[0];
                    dr = dx[0] * dx[0];
                    for (int d = 1; d < dim; d++) { dx[d] = xi[d] - xj[d]; dr = dr + dx[d] * dx[d]; }
                    r = sqrt(dr);
                    double eta = r / 0.75 / ((s->h[i] + s->h[j]) * 0.5);
                    double chi = 1 - r / cfg->couple_delta;
                    if (chi < 0 || chi > 1) chi = 0;
                    double fc = 0;
                    if (eta <= 0 || eta >= 2) fc = 0;
                    else if (eta <= 2.0 / 3) fc = 2.0 / 3;
                    else if (eta <= 1) fc = 2 * eta - 1.5 * eta * eta;
                    else if (eta < 2) fc = 0.5 * ((2 - eta) * (2 - eta));
                    double f = (double)0.01f * (s->c[i] * s->c[i]) * chi * fc;
                    for (int d = 0; d < dim; d++) acc[d] = acc[d] + f * dx[d] / (r * r);
                }
            }
        }
        for (int d = 0; d < dim; d++) dvdt[(int64_t)i * dim + d] = acc[d];
    }
}

/* artificialViscosity.f90:7-114 */
static void or_arti_visc(const sphgpu_config* cfg, or_state* s, const or_pairs* pr,
                         double* dvdt, double* dedt) {
    int dim = s->dim, nt = s->ntotal;
    double alpha = cfg->av_alpha, beta = cfg->av_beta, psi = cfg->av_psi;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nt; i++) {
        double acc[3] = {0, 0, 0}, de = 0;
        const double* xi = s->x + (int64_t)i * dim;
        const double* vi = s->v + (int64_t)i * dim;
        for (int64_t k = pr->off[i]; k < pr->off[i + 1]; k++) {
            int j = pr->nbr[k];
            const double* xj = s->x + (int64_t)j * dim;
            const double* vj = s->v + (int64_t)j * dim;
            double dx[3], dv[3], xv = 0, sx2 = 0;
            for (int d = 0; d < dim; d++) { dx[d] = xi[d] - xj[d]; dv[d] = vi[d] - vj[d]; }
            for (int d = 0; d < dim; d++) xv = xv + dx[d] * dv[d];
            if (xv < 0.0) {
                double hij = 0.5 * (s->h[i] + s->h[j]);
                double rhoij = 0.5 * (s->rho[i] + s->rho[j]);
                double cij = 0.5 * (s->c[i] + s->c[j]);
                for (int d = 0; d < dim; d++) sx2 = sx2 + dx[d] * dx[d];
                double ph = psi * hij;
                double phi = hij * xv / (sx2 + ph * ph);
                double PIij = (beta * phi - alpha * cij) * phi / rhoij;
                double aux[3], sd = 0;
                for (int d = 0; d < dim; d++) aux[d] = PIij * pr->dwdx[k * dim + d];
                for (int d = 0; d < dim; d++) acc[d] = acc[d] - s->mass[j] * aux[d];
                for (int d = 0; d < dim; d++) sd = sd + dv[d] * aux[d];
                de = de + 0.5 * s->mass[j] * sd;
            }
        }
        for (int d = 0; d < dim; d++) dvdt[(int64_t)i * dim + d] = acc[d];
        dedt[i] = de;
    }
}

/* artificialTerm.f90:7-98 (2D). */
static int or_artificial_stress(const sphgpu_config* cfg, or_state* s, const or_pairs* pr,
                                double* dvdt) {
    int dim = s->dim, nt = s->ntotal;
    if (dim != 2) return SPHGPU_E_BAD_CONFIG; /* artificialTerm.f90:37-39, :94-95 */
    double lambda = cfg->as_lambda;
    double* R = (double*)calloc((size_t)nt * 4, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nt; i++) {
        if (s->type[i] < 100) continue;
        const double* St = s->stress + (int64_t)i * 4;
        double aux = St[0] - St[3], theta, c, sn, ps[2], T[2];
        if (aux == 0) theta = 0.5 * PI_D;
        else theta = 0.5 * atan((St[1] + St[2]) / aux);
        sn = sin(theta);
        c = cos(theta);
        ps[0] = c * c * St[0] + sn * sn * St[3] + 2 * sn * c * St[1];
        ps[1] = sn * sn * St[0] + c * c * St[3] - 2 * sn * c * St[1];
        for (int d = 0; d < 2; d++)
            T[d] = ps[d] > 0 ? -lambda * ps[d] / (s->rho[i] * s->rho[i]) : 0;
        R[i * 4 + 0] = c * c * T[0] + sn * sn * T[1];
        R[i * 4 + 3] = sn * sn * T[0] + c * c * T[1];
        R[i * 4 + 1] = sn * c * (T[0] - T[1]);
        R[i * 4 + 2] = R[i * 4 + 1];
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nt; i++) {
        double acc[2] = {0, 0};
        for (int64_t k = pr->off[i]; k < pr->off[i + 1]; k++) {
            int j = pr->nbr[k];
            if (s->type[j] < 100) continue;
            double zero[2] = {0, 0}, wd, tmp[2];
            or_kernel(dim, cfg->skf, cfg->as_delta, zero, s->h[i], &wd, tmp);
            double F = pow(pr->w[k] / wd, cfg->as_exponent);
            static const double zero4[4] = {0, 0, 0, 0};
            const double* Rj = (j < nt) ? R + (int64_t)j * 4 : zero4;
            for (int d = 0; d < 2; d++)
                for (int e = 0; e < 2; e++)
                    acc[d] = acc[d] + s->mass[j] * (R[i * 4 + d * 2 + e] + Rj[d * 2 + e]) * F * pr->dwdx[k * 2 + e];
        }
        dvdt[(int64_t)i * 2 + 0] = acc[0];
        dvdt[(int64_t)i * 2 + 1] = acc[1];
    }
    free(R);
    return 0;
}

/* hsml.f90:8-38 (h_upgrade), real particles. */
static void or_h_upgrade(const sphgpu_config* cfg, or_state* s) {
    int dim = s->dim, nt = s->ntotal;
    if (cfg->sle == 1) {
        double factor = 2.0;
        for (int i = 0; i < nt; i++) s->h[i] = factor * pow(s->mass[i] / s->rho[i], 1.0 / dim);
    } else if (cfg->sle == 2) {
        for (int i = 0; i < nt; i++) {
            double dh = (s->h[i] / dim) * s->divv[i];
            s->h[i] = s->h[i] + cfg->dt * dh;
            if (s->h[i] <= 0) s->h[i] = s->h[i] - cfg->dt * dh;
        }
    }
}

/* singleStep.f90:47-144. Fills s->dvdt, s->drhodt (continuity only), s->dedt,
 * s->dsdt and updates rho (summation), p, eta, stress, divx, divv. */
static int or_single_step(const sphgpu_config* cfg, or_state* s, or_pairs* keep) {
    int dim = s->dim, nt = s->ntotal;
    or_pairs pr;
    int rc = or_bggs(dim, cfg->skf, nt, s->n, s->x, s->h, s->state, &pr, NULL);
    if (rc) return rc;
    double* indvdt = (double*)calloc((size_t)nt * dim + 1, sizeof(double));
    double* exdvdt = (double*)calloc((size_t)nt * dim + 1, sizeof(double));
    double* avdvdt = (double*)calloc((size_t)nt * dim + 1, sizeof(double));
    double* asdvdt = (double*)calloc((size_t)nt * dim + 1, sizeof(double));
    double* indedt = (double*)calloc((size_t)nt + 1, sizeof(double));
    double* avdedt = (double*)calloc((size_t)nt + 1, sizeof(double));
    if (cfg->kernel_correction) or_kgc(s, &pr);
    /* singleStep.f90:90-105: SPH = 3 takes the Riemann rate whatever the
     * summation switch says */
    if (cfg->pa_sph == 3) or_con_density_riemann(s, &pr, s->drhodt);
    else if (cfg->summation_density) or_sum_density(cfg, s, &pr);
    else or_con_density(s, &pr, s->drhodt);
    or_divergence(s, &pr);
    or_internal_force(cfg, s, &pr, indvdt, indedt);
    if (cfg->external_force) or_external_force(cfg, s, &pr, exdvdt);
    if (cfg->artificial_viscosity) or_arti_visc(cfg, s, &pr, avdvdt, avdedt);
    if (cfg->artificial_stress) rc = or_artificial_stress(cfg, s, &pr, asdvdt);
    or_h_upgrade(cfg, s);   /* singleStep.f90:136 */
    for (int i = 0; i < nt; i++) {
        for (int d = 0; d < dim; d++) {
            int64_t q = (int64_t)i * dim + d;
            s->dvdt[q] = indvdt[q] + avdvdt[q] + exdvdt[q] + asdvdt[q];
        }
        s->dedt[i] = indedt[i] + avdedt[i];
        if (s->boundary && s->boundary[i] == 1) /* boundaryCondition.f90:32-45 */
            for (int d = 0; d < dim; d++) s->dvdt[(int64_t)i * dim + d] = 0;
    }
    free(indvdt); free(exdvdt); free(avdvdt); free(asdvdt); free(indedt); free(avdedt);
    if (keep) *keep = pr; else or_pairs_free(&pr);
    return rc;
}

/* Shear update of leap_frog by P%Type (timeIntegration.f90:149-166,
 * :190-205, :225-240). mode 0: predict, 1: first step, 2: correct. */
static void or_shear_update(const sphgpu_config* cfg, or_state* s, int i, int mode) {
    int dim = s->dim;
    const sphgpu_material* m = or_mat(cfg, s->type[i]);
    if (s->type[i] <= 100 || !m) return;
    double* S = s->shear + (int64_t)i * dim * dim;
    double* Sp = s->shear_prev + (int64_t)i * dim * dim;
    const double* D = s->dsdt + (int64_t)i * dim * dim;
    int nn = dim * dim;
    double dt = cfg->dt;
    if (mode == 0) {
        for (int q = 0; q < nn; q++) Sp[q] = S[q];
        for (int q = 0; q < nn; q++) S[q] = S[q] + (dt / 2) * D[q];
    } else if (mode == 1) {
        for (int q = 0; q < nn; q++) S[q] = S[q] + (dt / 2) * D[q];
    } else {
        for (int q = 0; q < nn; q++) S[q] = Sp[q] + dt * D[q];
    }
    if (m->sigma_y > 0) {
        double J2 = 0;
        for (int e = 0; e < dim; e++)
            for (int d = 0; d < dim; d++) J2 = J2 + S[d * dim + e] * S[d * dim + e];
        double f = sqrt(((m->sigma_y * m->sigma_y) / 3) / J2);
        double sc = 1.0 < f ? 1.0 : f;
        for (int q = 0; q < nn; q++) S[q] = S[q] * sc;
    }
}

/* leap_frog, timeIntegration.f90:137-252, for steps first_step ..
 * first_step+nsteps-1 (step 1 is the reference's first step). */
static int or_leap_frog(const sphgpu_config* cfg, or_state* s, int64_t first_step, int nsteps) {
    int dim = s->dim, nt = s->ntotal;
    double dt = cfg->dt;
    for (int it = 0; it < nsteps; it++) {
        int64_t step = first_step + it;
        if (step != 1) {
            for (int i = 0; i < nt; i++) {
                s->e_prev[i] = s->e[i];
                s->e[i] = s->e[i] + (dt / 2) * s->dedt[i];
                if (s->e[i] < 0) s->e[i] = 0;
                if (!cfg->summation_density) {
                    s->rho_prev[i] = s->rho[i];
                    s->rho[i] = s->rho[i] + (dt / 2) * s->drhodt[i];
                }
                for (int d = 0; d < dim; d++) {
                    int64_t q = (int64_t)i * dim + d;
                    s->v_prev[q] = s->v[q];
                    s->v[q] = s->v[q] + (dt / 2) * s->dvdt[q];
                }
                or_shear_update(cfg, s, i, 0);
            }
        }
        int rc = or_single_step(cfg, s, NULL);
        if (rc) return rc;
        if (step == 1) {
            for (int i = 0; i < nt; i++) {
                s->e[i] = s->e[i] + (dt / 2) * s->dedt[i];
                if (s->e[i] < 0) s->e[i] = 0;
                if (!cfg->summation_density) s->rho[i] = s->rho[i] + (dt / 2) * s->drhodt[i];
                for (int d = 0; d < dim; d++) {
                    int64_t q = (int64_t)i * dim + d;
                    s->v[q] = s->v[q] + (dt / 2) * s->dvdt[q];
                    s->disp[q] = dt * s->v[q];
                    s->x[q] = s->x[q] + s->disp[q];
                }
                or_shear_update(cfg, s, i, 1);
            }
        } else {
            for (int i = 0; i < nt; i++) {
                s->e[i] = s->e_prev[i] + dt * s->dedt[i];
                if (s->e[i] < 0) s->e[i] = 0;
                if (!cfg->summation_density) s->rho[i] = s->rho_prev[i] + dt * s->drhodt[i];
                for (int d = 0; d < dim; d++) {
                    int64_t q = (int64_t)i * dim + d;
                    s->v[q] = s->v_prev[q] + dt * s->dvdt[q];
                    s->disp[q] = s->disp[q] + dt * s->v[q];
                    s->x[q] = s->x[q] + dt * s->v[q];
                }
                or_shear_update(cfg, s, i, 2);
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------- public API */

/* Flat view of the particle state for the exported calls. All arrays hold
 * n = ntotal + ndummy particles; integrator arrays (shear_prev, v_prev,
 * rho_prev, e_prev) and derivative arrays hold at least ntotal. */
typedef struct oracle_arrays {
    int32_t dim, ntotal, n, pad;
    double *x, *v, *disp, *mass, *rho, *p, *e, *c, *h, *eta, *divx, *divv, *stress;
    const int32_t *type, *state, *boundary;
    double *shear, *shear_prev, *dsdt, *v_prev, *rho_prev, *e_prev;
    double *dvdt, *drhodt, *dedt;
} oracle_arrays;

static void or_bind(const oracle_arrays* a, or_state* s) {
    s->dim = a->dim; s->ntotal = a->ntotal; s->n = a->n;
    s->x = a->x; s->v = a->v; s->disp = a->disp; s->mass = a->mass; s->rho = a->rho;
    s->p = a->p; s->e = a->e; s->c = a->c; s->h = a->h; s->eta = a->eta;
    s->divx = a->divx; s->divv = a->divv; s->stress = a->stress;
    s->type = a->type; s->state = a->state; s->boundary = a->boundary;
    s->shear = a->shear; s->shear_prev = a->shear_prev; s->dsdt = a->dsdt;
    s->v_prev = a->v_prev; s->rho_prev = a->rho_prev; s->e_prev = a->e_prev;
    s->dvdt = a->dvdt; s->drhodt = a->drhodt; s->dedt = a->dedt;
}

int oracle_abi_size(void) { return (int)sizeof(oracle_arrays); }
int oracle_config_size(void) { return (int)sizeof(sphgpu_config); }

int oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* One single_step evaluation (derivatives at the current state). */
int oracle_derivatives(const sphgpu_config* cfg, const oracle_arrays* a) {
    or_state s;
    or_bind(a, &s);
    return or_single_step(cfg, &s, NULL);
}

/* The external force alone at the current state (singleStep.f90:84 then
 * :116): BGGS of the state, then external_force into dvdt (ntotal x dim). */
int oracle_external_force(const sphgpu_config* cfg, const oracle_arrays* a, double* dvdt) {
    or_state s;
    or_bind(a, &s);
    or_pairs pr;
    int rc = or_bggs(s.dim, cfg->skf, s.ntotal, s.n, s.x, s.h, s.state, &pr, NULL);
    if (rc) return rc;
    or_external_force(cfg, &s, &pr, dvdt);
    or_pairs_free(&pr);
    return 0;
}

/* nsteps of leap_frog starting at first_step (1 = fresh start). */
int oracle_run(const sphgpu_config* cfg, const oracle_arrays* a, int64_t first_step, int nsteps) {
    or_state s;
    or_bind(a, &s);
    return or_leap_frog(cfg, &s, first_step, nsteps);
}
