/*
 * sanitize_check.c - TEST INFRASTRUCTURE ONLY: the oracle under
 * AddressSanitizer and UndefinedBehaviorSanitizer.
 *
 * Compiled together with sph_oracle.c (included below) by
 * tests/test_oracle_sanitize.py:
 *   gcc -g -O1 -fsanitize=address,undefined -fno-sanitize-recover=all \
 *       -ffp-contract=off oracle/sanitize_check.c -lm
 * It runs BGGS and leap-frog steps of the oracle on small seeded particle
 * sets, in 2D and 3D, through every switch the oracle implements (cubic,
 * Gauss and quintic kernels, continuity and summation density with Shepard
 * normalisation, physical viscosity, the wall / penalty / coupling forces,
 * artificial viscosity and stress, kernel-gradient correction, SPH = 1 and 3,
 * smoothing-length evolution 1 and 2, State == -1 particles, fixed-boundary
 * particles, the J2 cap), and exits 0 when every call returned what it
 * should. Any out-of-bounds access, leak, use after free, signed overflow or
 * other undefined behaviour aborts the run with the sanitizer's report.
 */
#include "sph_oracle.c"

#include <stdio.h>

static uint64_t rng_state = 20261209ull;
static double urand(void) { /* splitmix64 -> [0, 1) */
    uint64_t z = (rng_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

typedef struct {
    int dim, nt, n;
    double *x, *v, *disp, *mass, *rho, *p, *e, *c, *h, *eta, *divx, *divv, *stress;
    int32_t *type, *state, *boundary;
    double *shear, *shear_prev, *dsdt, *v_prev, *rho_prev, *e_prev, *dvdt, *drhodt, *dedt;
} set_t;

static double* zd(size_t k) { return (double*)calloc(k ? k : 1, sizeof(double)); }
static int32_t* zi(size_t k) { return (int32_t*)calloc(k ? k : 1, sizeof(int32_t)); }

/* A jittered block of fluid (types 2 and 3; with all_solid, of the solid
 * type), a solid strip (type 102 or 104) and a layer of fixed particles
 * (type -2, sound speed dummy_c). */
static int all_solid = 0;
static double dummy_c = 0.0;
static set_t make_set(int dim, int with_solid, int solid_type) {
    set_t s;
    memset(&s, 0, sizeof(s));
    s.dim = dim;
    const int nx = 10, ny = 8, nz = dim == 3 ? 4 : 1;
    const int nf = nx * ny * nz, ns = with_solid ? 2 * ny * nz : 0, nw = nx * 2 * nz;
    s.nt = nf + ns;
    s.n = s.nt + nw;
    size_t n = (size_t)s.n, d = (size_t)dim, dd = d * d;
    s.x = zd(n * d); s.v = zd(n * d); s.disp = zd(n * d); s.mass = zd(n); s.rho = zd(n);
    s.p = zd(n); s.e = zd(n); s.c = zd(n); s.h = zd(n); s.eta = zd(n); s.divx = zd(n);
    s.divv = zd(n); s.stress = zd(n * dd); s.type = zi(n); s.state = zi(n); s.boundary = zi(n);
    s.shear = zd(n * dd); s.shear_prev = zd(n * dd); s.dsdt = zd(n * dd); s.v_prev = zd(n * d);
    s.rho_prev = zd(n); s.e_prev = zd(n); s.dvdt = zd(n * d); s.drhodt = zd(n); s.dedt = zd(n);
    const double dx = 0.01;
    int k = 0;
    for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++, k++) {
                s.x[k * d + 0] = (ix + 0.5 + 0.3 * (urand() - 0.5)) * dx;
                s.x[k * d + 1] = (iy + 2.5 + 0.3 * (urand() - 0.5)) * dx;
                if (dim == 3) s.x[k * d + 2] = (iz + 0.5 + 0.3 * (urand() - 0.5)) * dx;
                s.type[k] = all_solid ? solid_type : ix < nx / 2 ? 2 : 3;
                s.rho[k] = (all_solid ? 1100.0 : 1000.0) + 5.0 * (urand() - 0.5);
                s.c[k] = 20.0;
                s.h[k] = 1.2 * dx * (1.0 + 0.1 * urand());
                for (int q = 0; q < dim; q++) s.v[k * d + q] = 0.5 * (urand() - 0.5);
            }
    for (int iz = 0; iz < nz && with_solid; iz++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < 2; ix++, k++) {
                s.x[k * d + 0] = (nx + ix + 0.5) * dx * 0.98;
                s.x[k * d + 1] = (iy + 2.5) * dx;
                if (dim == 3) s.x[k * d + 2] = (iz + 0.5) * dx;
                s.type[k] = solid_type;
                s.rho[k] = solid_type == 104 ? 2700.0 : 1100.0;
                s.c[k] = 130.0;
                s.h[k] = 1.2 * dx;
                s.boundary[k] = iy == 0;
                for (int q = 0; q < dim; q++) s.v[k * d + q] = 2.0 * (urand() - 0.5);
            }
    for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < 2; iy++)
            for (int ix = 0; ix < nx; ix++, k++) {
                s.x[k * d + 0] = (ix + 0.5) * dx;
                s.x[k * d + 1] = (iy + 0.5) * dx;
                if (dim == 3) s.x[k * d + 2] = (iz + 0.5) * dx;
                s.type[k] = -2;
                s.rho[k] = 1000.0;
                s.c[k] = dummy_c;
                s.h[k] = 1.2 * dx;
            }
    for (int i = 0; i < s.n; i++) s.mass[i] = s.rho[i] * pow(dx, dim);
    s.state[3] = -1; /* a particle out of the search */
    return s;
}

static void free_set(set_t* s) {
    double* dp[] = {s->x, s->v, s->disp, s->mass, s->rho, s->p, s->e, s->c, s->h, s->eta, s->divx,
                    s->divv, s->stress, s->shear, s->shear_prev, s->dsdt, s->v_prev, s->rho_prev,
                    s->e_prev, s->dvdt, s->drhodt, s->dedt};
    for (size_t i = 0; i < sizeof(dp) / sizeof(dp[0]); i++) free(dp[i]);
    free(s->type); free(s->state); free(s->boundary);
}

static oracle_arrays bind(set_t* s) {
    oracle_arrays a;
    memset(&a, 0, sizeof(a));
    a.dim = s->dim; a.ntotal = s->nt; a.n = s->n;
    a.x = s->x; a.v = s->v; a.disp = s->disp; a.mass = s->mass; a.rho = s->rho; a.p = s->p;
    a.e = s->e; a.c = s->c; a.h = s->h; a.eta = s->eta; a.divx = s->divx; a.divv = s->divv;
    a.stress = s->stress; a.type = s->type; a.state = s->state; a.boundary = s->boundary;
    a.shear = s->shear; a.shear_prev = s->shear_prev; a.dsdt = s->dsdt; a.v_prev = s->v_prev;
    a.rho_prev = s->rho_prev; a.e_prev = s->e_prev; a.dvdt = s->dvdt; a.drhodt = s->drhodt;
    a.dedt = s->dedt;
    return a;
}

static sphgpu_config base(int dim) {
    sphgpu_config c;
    memset(&c, 0, sizeof(c));
    c.dim = dim; c.skf = 1; c.pa_sph = 2; c.gravity = 1; c.gravity_g = -9.81;
    c.external_force = 1; c.artificial_viscosity = 1;
    c.av_alpha = 1; c.av_beta = 1; c.av_psi = 0.1;
    c.lj_factor = 1e-2; c.lj_r0 = 0.004; c.lj_p1 = 12; c.lj_p2 = 4;
    c.pen_factor = 1e2; c.pen_n1 = 6; c.pen_n2 = 4; c.couple_delta = 0.012;
    c.as_lambda = 0.2; c.as_delta = 0.01; c.as_exponent = 4;
    c.dt = 1e-6;
    c.n_materials = 4;
    c.mat[0] = (sphgpu_material){2, SPHGPU_EOS_TAIT, 1000.0, 20.0 * 20.0 * 1000.0 / 7.0, 7.0, 1e-3, 0.0};
    c.mat[1] = (sphgpu_material){3, SPHGPU_EOS_TAIT, 1000.0, 20.0 * 20.0 * 1000.0 / 7.0, 7.0, 1e-3, 0.0};
    c.mat[2] = (sphgpu_material){102, SPHGPU_EOS_LINEAR, 1100.0, 2e7 / 1100.0, 1.0, 4.27e6, 0.0};
    c.mat[3] = (sphgpu_material){104, SPHGPU_EOS_LINEAR, 2700.0, 7.235e10 / 2700.0, 1.0, 2.77e10, 3e8};
    return c;
}

static int failures = 0;
static void expect(int cond, const char* what) {
    if (!cond) {
        fprintf(stderr, "FAILED: %s\n", what);
        failures++;
    }
}

static void run_case(const char* name, int dim, int with_solid, int solid_type, sphgpu_config cfg,
                     int steps) {
    set_t s = make_set(dim, with_solid, solid_type);
    oracle_arrays a = bind(&s);
    int rc = oracle_derivatives(&cfg, &a);
    if (rc) fprintf(stderr, "%s: derivatives rc %d\n", name, rc);
    expect(rc == 0, name);
    rc = oracle_run(&cfg, &a, 1, steps);
    if (rc) fprintf(stderr, "%s: run rc %d\n", name, rc);
    expect(rc == 0, name);
    int finite = 1;
    for (int i = 0; i < s.nt * dim; i++) finite &= isfinite(s.x[i]) && isfinite(s.v[i]);
    if (!finite) fprintf(stderr, "%s: non-finite state\n", name);
    expect(finite, name);
    free_set(&s);
    printf("ok %s\n", name);
}

int main(void) {
#ifdef SANITIZE_PLANT
    /* a planted read one element past an array: the test that builds this
     * expects AddressSanitizer to stop here */
    {
        set_t s = make_set(2, 0, 102);
        volatile double past = s.x[(size_t)s.n * 2];
        printf("read past the end: %g\n", past);
        free_set(&s);
    }
#endif
    for (int dim = 2; dim <= 3; dim++) {
        /* BGGS alone, every kernel, with capacity too small then right */
        for (int skf = 1; skf <= 3; skf++) {
            set_t s = make_set(dim, 1, 102);
            int64_t* off = (int64_t*)calloc((size_t)s.nt + 1, sizeof(int64_t));
            int64_t need = 0;
            double grid[9];
            int rc = oracle_bggs(dim, skf, s.nt, s.n, s.x, s.h, s.state, 0, off, NULL, NULL, NULL,
                                 &need, grid);
            expect(rc == SPHGPU_E_PAIR_CAPACITY || need == 0, "bggs sizing");
            int32_t* nbr = (int32_t*)calloc((size_t)need + 1, sizeof(int32_t));
            double* w = (double*)calloc((size_t)need + 1, sizeof(double));
            double* dw = (double*)calloc((size_t)(need + 1) * dim, sizeof(double));
            rc = oracle_bggs(dim, skf, s.nt, s.n, s.x, s.h, s.state, need, off, nbr, w, dw, &need, grid);
            expect(rc == 0 && off[s.nt] == need, "bggs");
            free(off); free(nbr); free(w); free(dw);
            free_set(&s);
        }
        sphgpu_config c = base(dim);
        run_case(dim == 2 ? "2D fluid, continuity" : "3D fluid, continuity", dim, 0, 102, c, 5);
        c = base(dim); c.summation_density = 1; c.normalize_density = 1;
        run_case("summation density + Shepard", dim, 1, 102, c, 5);
        c = base(dim); c.viscosity = 1; c.fluid_visc_skip_solid = 1;
        run_case("fluid and solid, viscosity", dim, 1, 102, c, 5);
        /* kernel-gradient correction solves A grad W = grad W with A = 0
         * for a fluid particle (kernelGradientCorrection.f90:20-45): the
         * reference uses it on all-solid bodies */
        all_solid = 1;
        c = base(dim); c.viscosity = 1; c.kernel_correction = 1;
        run_case("solid, viscosity, kernel correction", dim, 1, 102, c, 5);
        all_solid = 0;
        c = base(dim); c.viscosity = 1;
        run_case("type 104 with the J2 cap", dim, 1, 104, c, 5);
        for (int skf = 2; skf <= 3; skf++) {
            c = base(dim); c.skf = skf;
            run_case("Gauss / quintic", dim, 1, 102, c, 3);
        }
        c = base(dim); c.pa_sph = 1;
        run_case("SPH = 1", dim, 1, 102, c, 3);
        /* the Riemann forms divide by rho_i c_i + rho_j c_j */
        dummy_c = 20.0;
        c = base(dim); c.pa_sph = 3;
        run_case("SPH = 3", dim, 0, 102, c, 3);
        dummy_c = 0.0;
        for (int sle = 1; sle <= 2; sle++) {
            c = base(dim); c.sle = sle;
            run_case("smoothing-length evolution", dim, 0, 102, c, 3);
        }
        if (dim == 2) {
            c = base(dim); c.viscosity = 1; c.artificial_stress = 1;
            run_case("artificial stress", dim, 1, 102, c, 5);
        } else {
            c = base(dim); c.artificial_stress = 1;
            set_t s = make_set(dim, 1, 102);
            oracle_arrays a = bind(&s);
            expect(oracle_derivatives(&c, &a) == SPHGPU_E_BAD_CONFIG, "artificial stress in 3D refused");
            free_set(&s);
        }
    }
    if (failures) return 1;
    printf("sanitized oracle: all cases ok\n");
    return 0;
}
