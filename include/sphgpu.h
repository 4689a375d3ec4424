/*
 * sphgpu.h - C ABI of the MI355X SPH hot path (libsphgpu.so).
 *
 * The library replaces the per-step subtree of the reference solver:
 * the external subroutine single_step(ntotal, ndummy, nbuffer, Particles,
 * Delta, aver_v, Shear, dSdt) (src/TimeIntegration/singleStep.f90:1-211,
 * interface src/TimeIntegration/timeIntegration.f90:7-19) and, inside it, the
 * src/NNPS and src/Physic calls listed at singleStep.f90:58-144, plus the
 * per-particle update loops of leap_frog (timeIntegration.f90:137-252).
 *
 * Conventions
 *  - Plain pointers and sizes only. Arrays are particle-major: a vector field
 *    is x[i*dim + d], a tensor field is s[(i*dim + d)*dim + e]. That is the
 *    memory order of Fortran x(dim, N) and Stress(dim, dim, N) with d and e
 *    swapped for tensors (Fortran S(d,e,i) sits at s[(i*dim + e)*dim + d]);
 *    the Fortran shim passes c_loc of contiguous arrays and transposes
 *    tensors.
 *  - Particle i < n_real is a real (integrated) particle, i >= n_real a fixed
 *    boundary ("dummy") particle, in the order the reference keeps them
 *    (Particles(1:ntotal), then ntotal+1:ntotal+ndummy).
 *  - Every call returns 0 or a negative sphgpu_status; sphgpu_last_error()
 *    gives the message. The Fortran shim turns a nonzero code into
 *    `error stop`, as the reference does (asymmetricSearch.f90:38-39).
 *  - One context = one HIP stream = one host thread. Calls are not
 *    thread-safe on the same context.
 */
#ifndef SPHGPU_H
#define SPHGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPHGPU_ABI_VERSION 2
#define SPHGPU_MAX_MATERIALS 8

typedef enum sphgpu_status {
    SPHGPU_OK = 0,
    SPHGPU_E_OUT_OF_RANGE = -1,   /* particle outside the grid (asymmetricSearch.f90:38-39) */
    SPHGPU_E_NONFINITE = -2,      /* NaN/Inf position, density or velocity */
    SPHGPU_E_PAIR_CAPACITY = -3,  /* pair export buffer too small; or, from a step on
                                   * the general path, a particle outgrew the exact
                                   * neighbour list planned at upload (1.5x the largest
                                   * count + 8; the reference grows its list): that
                                   * step's sums are incomplete, upload the last good
                                   * state again. A particle that outgrows its candidate
                                   * list (fast path, single or batched, and general
                                   * path) walks its cell window instead and stays
                                   * exact; a batched run rebuilds every step while it
                                   * has one. */
    SPHGPU_E_BAD_CONFIG = -4,     /* unsupported switch combination */
    SPHGPU_E_HIP = -5,            /* HIP runtime error */
    SPHGPU_E_ARG = -6,            /* bad argument (NULL, size) */
    SPHGPU_E_CAPACITY = -7        /* more particles than the context holds */
} sphgpu_status;

/* Arithmetic of the pair kernels. Positions, velocities, the neighbour test
 * and the integrator stay FP64 in both modes. MIXED on the fluid set of the
 * fast path also stores density as an FP32 deviation rho - rho0 and the pair
 * inputs as FP32 records; outside that set (solids, viscosity, kernel
 * correction, artificial stress, summation density) the state stays FP64 and
 * only the pair sums run in FP32. */
typedef enum sphgpu_precision {
    SPHGPU_FP64 = 0,
    SPHGPU_MIXED = 1              /* FP32 pair physics */
} sphgpu_precision;

typedef enum sphgpu_eos {
    SPHGPU_EOS_NONE = 0,          /* pressure left unchanged */
    SPHGPU_EOS_TAIT = 1,          /* p = B((rho/rho0)^gamma - 1)   eos.f90:23-53 */
    SPHGPU_EOS_LINEAR = 2,        /* p = K (rho - rho0)            eos.f90:149-157, 219-248 */
    SPHGPU_EOS_ISOTHERMAL = 3     /* p = K rho                     eos.f90:57-64 */
} sphgpu_eos;

/* Material of the particles whose |Type| equals `type`. The reference
 * hard-codes these per type and per case nick (eos.f90, viscosity.f90,
 * timeIntegration.f90:149-166); here they are explicit. */
typedef struct sphgpu_material {
    int32_t type;        /* |Type| this entry applies to */
    int32_t eos;         /* sphgpu_eos */
    double rho0;         /* reference density */
    double K;            /* Tait: B = c^2 rho0 / gamma; LINEAR: K/rho0; ISOTHERMAL: c^2 */
    double gamma;        /* Tait exponent */
    double eta;          /* viscosity.f90:14-33: dynamic viscosity, or shear modulus G for solids */
    double sigma_y;      /* J2 cap (timeIntegration.f90:149-166); 0 = none */
} sphgpu_material;

typedef enum sphgpu_gravity_mode {
    SPHGPU_GRAVITY_ALL = 0,       /* externalForce.f90:54-57 (default), :49-53 */
    SPHGPU_GRAVITY_SKIP_SOLID = 1 /* externalForce.f90:42-48 (Type > 100 skipped) */
} sphgpu_gravity_mode;

/* Switches and constants of one run. Field names follow the reference's
 * Config/Field globals (src/Initialize/controlDict.f90:16-74) and the
 * constants its routines select by case nick. */
typedef struct sphgpu_config {
    int32_t dim;                  /* Field%Dim: 2 or 3 */
    int32_t skf;                  /* Config%skf: 1 cubic spline, 2 Gauss, 3 quintic */
    int32_t pa_sph;               /* Config%pa_sph: 2 (internalForce.f90:228-274) */
    int32_t summation_density;    /* Config%sum_density_w: density.f90:10-80, else :83-107 */
    int32_t normalize_density;    /* Config%norm_dens_w (Shepard, density.f90:25-76) */
    int32_t viscosity;            /* Config%viscosity_w */
    int32_t gravity;              /* Config%gravity_w */
    int32_t gravity_mode;         /* sphgpu_gravity_mode */
    int32_t external_force;       /* Config%ex_force_w (pair forces of externalForce.f90:103-174) */
    int32_t artificial_viscosity; /* Config%arti_visc_w */
    int32_t artificial_stress;    /* Config%arti_stress_w (2D only, artificialTerm.f90:37-96) */
    int32_t kernel_correction;    /* Config%kernel_correciton_w (kernelGradientCorrection.f90) */
    int32_t precision;            /* sphgpu_precision */
    int32_t n_materials;
    double gravity_g;             /* externalForce.f90:29: -9.81 on the last axis */
    double av_alpha, av_beta, av_psi;            /* artificialViscosity.f90:28-79 */
    double lj_factor, lj_r0, lj_p1, lj_p2;       /* externalForce.f90:62-101 (factor_s, r0, p1, p2) */
    double pen_factor, pen_n1, pen_n2;           /* externalForce.f90:68-70 (factor_p, n1, n2) */
    double couple_delta;                         /* externalForce.f90:72-97 (delta) */
    double as_lambda, as_delta, as_exponent;     /* artificialTerm.f90:17-32 */
    double dt;                                   /* Config%delta_t */
    sphgpu_material mat[SPHGPU_MAX_MATERIALS];
    /* No counterpart in the reference. 1: the physical-viscosity pair term
     * of a fluid particle (internalForce.f90:236-244) skips solid
     * neighbours, whose eta is their shear modulus G (viscosity.f90:24).
     * With 0 (the reference), a fluid next to an elastic solid takes
     * G * strain rate as a viscous stress, which no explicit time step can
     * integrate stably (diffusion number G dt / (rho h^2) >> 1). */
    int32_t fluid_visc_skip_solid;
    /* Config%sle, smoothing-length evolution after the force pass
     * (singleStep.f90:136, hsml.f90:8-38), real particles only: 0 keep h,
     * 1 h = 2 (m/rho)^(1/dim), 2 h += dt (h/dim) divergenceVelocity. The
     * search grid and the neighbour capacity are planned for h up to 1.5x
     * the larger of its uploaded value and (sle 1) its first update; beyond
     * that a step fails with SPHGPU_E_OUT_OF_RANGE (upload again). */
    int32_t sle;
} sphgpu_config;

/* Field selectors for sphgpu_download / sphgpu_upload_field. */
typedef enum sphgpu_field {
    SPHGPU_F_X = 0,       /* dim per particle */
    SPHGPU_F_V = 1,       /* dim */
    SPHGPU_F_RHO = 2,
    SPHGPU_F_P = 3,
    SPHGPU_F_E = 4,
    SPHGPU_F_MASS = 5,
    SPHGPU_F_H = 6,
    SPHGPU_F_C = 7,
    SPHGPU_F_DISP = 8,    /* dim */
    SPHGPU_F_STRESS = 9,  /* dim*dim, Stress = Shear - p I (internalForce.f90:175-180) */
    SPHGPU_F_SHEAR = 10,  /* dim*dim, deviatoric stress state of the integrator */
    SPHGPU_F_DVDT = 11,   /* dim, Delta%Velocity */
    SPHGPU_F_DRHODT = 12, /* Delta%Density */
    SPHGPU_F_DEDT = 13,   /* Delta%Energy */
    SPHGPU_F_DSDT = 14,   /* dim*dim */
    SPHGPU_F_DIVX = 15,   /* divergencePosition */
    SPHGPU_F_DIVV = 16,   /* divergenceVelocity */
    SPHGPU_F_ETA = 17,    /* Viscosity */
    SPHGPU_F_NFIELDS = 18
} sphgpu_field;

typedef struct sphgpu_ctx sphgpu_ctx;

/* Library / ABI version, for a loader's sanity check. */
int sphgpu_abi_version(void);

/* Create a context on the current HIP device holding up to `capacity`
 * particles (real + dummy). Replaces allocateParticleList
 * (src/Initialize/SPH.f90:126-136) for the device copy of the state. */
int sphgpu_create(const sphgpu_config* cfg, int64_t capacity, sphgpu_ctx** out);

/* Copy a particle set to the device (host keeps ownership). Any pointer
 * except x, mass, rho, h, type may be NULL (zero). Resets the integrator
 * (the next step is step 1, timeIntegration.f90:178-212). `shear` is the
 * initial deviatoric stress (NULL = 0). */
int sphgpu_upload(sphgpu_ctx* ctx, int64_t n_real, int64_t n_dummy,
                  const double* x, const double* v, const double* mass,
                  const double* rho, const double* p, const double* e,
                  const double* c, const double* h, const double* shear,
                  const int32_t* type, const int32_t* state,
                  const int32_t* boundary);

/* One evaluation of the derivatives at the resident state: BGGS neighbour
 * set, density, EOS, internal/external/artificial forces and the fixed
 * boundary, in the order of singleStep.f90:58-144. Like the reference it
 * updates rho (summation density), p, stress, viscosity and the
 * divergences in the resident state. Results stay on the device (read them
 * with sphgpu_download). */
int sphgpu_derivatives(sphgpu_ctx* ctx);

/* Advance nsteps leap-frog steps on the device (timeIntegration.f90:
 * 137-252); the host keeps output cadence. The first step after an upload
 * is the reference's step 1. */
int sphgpu_advance(sphgpu_ctx* ctx, int32_t nsteps);

/* Copy one field of all n_real + n_dummy particles (in upload order) to
 * host memory `dst` (doubles; count = n * components of the field). */
int sphgpu_download(sphgpu_ctx* ctx, int32_t field, double* dst);

/* Copy one field from host memory `src` (upload order, as sphgpu_download
 * writes it) into the resident state. Supported: SPHGPU_F_DIVX and
 * SPHGPU_F_DIVV, the divergences a step leaves in the particles
 * (Particles%divergencePosition / divergenceVelocity, divergence.f90:6-29).
 * A caller that uploads the state on every step (the drop-in single_step)
 * passes the last step's divergencePosition this way: summation density
 * with Shepard normalisation normalises only particles whose previous
 * divergencePosition is below 1.5 (density.f90:25-37, :67-76). Other fields
 * return SPHGPU_E_ARG (upload them with sphgpu_upload). */
int sphgpu_upload_field(sphgpu_ctx* ctx, int32_t field, const double* src);

/* Neighbour pairs of the current state, as BGGS builds them
 * (asymmetricSearch.f90:160-257): for each real particle i, offsets[i] ..
 * offsets[i+1]-1 index nbr[], the neighbour ids in ascending order.
 * offsets has n_real+1 entries. If cap is too small, returns
 * SPHGPU_E_PAIR_CAPACITY and sets *needed. */
int sphgpu_export_pairs(sphgpu_ctx* ctx, int64_t* offsets, int32_t* nbr,
                        int64_t cap, int64_t* needed);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (serving loops that forbid it, harnesses that run under an
 * allocation ban): the pair set of sphgpu_export_pairs for a particle state
 * held in the caller's own device memory, with no context and no
 * allocation. All inputs and outputs are DEVICE pointers: x (n * dim,
 * particle-major), h (n), state (n, or NULL: all 0); offsets (n_real + 1)
 * and nbr (cap ids) receive the CSR result, ids ascending per target;
 * *needed (host) receives the pair count. Particles i < n_real are the
 * targets, all n are sources; skf picks the reach (2h cubic, 3h Gauss and
 * quintic). Work runs on `stream` (a hipStream_t, NULL for the null
 * stream), which the call synchronises before it returns.
 * Limit: n <= 1048576 particles, and a reference grid of at most 1114112
 * cells (SPHGPU_E_CAPACITY beyond either). The library's scratch for this
 * call is one static device array of about 18 MB per GPU, held until the
 * process ends and shared by every call on that GPU: calls that overlap on
 * one GPU race and must be serialised by the caller. cap too small:
 * SPHGPU_E_PAIR_CAPACITY with *needed set and offsets written. Failure
 * message: sphgpu_last_error(NULL). */
int sphgpu_pairs_noalloc(int32_t dim, int32_t skf, int64_t n_real, int64_t n, const double* x,
                         const double* h, const int32_t* state, int64_t* offsets, int32_t* nbr,
                         int64_t cap, int64_t* needed, void* stream);

/* Neighbour statistics of the last neighbour search, as print_statistics
 * reports them (src/NNPS/nnps.f90:377-416): out[0] total pairs, out[1] max
 * and out[2] min neighbours per real particle, out[3] real particles with
 * no neighbour. */
int sphgpu_neighbour_stats(sphgpu_ctx* ctx, int64_t* out4);

/* Geometry of the reference background grid for the current state:
 * out[0..2] minCoor, out[3..5] maxCoor, out[6..8] cellNum (as doubles). */
int sphgpu_grid_info(sphgpu_ctx* ctx, double* out9);

/* Milliseconds spent in the last sphgpu_advance, measured with HIP events
 * on the context's stream, and the average duration of the last call's
 * force kernel launches. */
int sphgpu_timing(sphgpu_ctx* ctx, double* advance_ms, double* force_kernel_ms,
                  int64_t* force_launches);

/* Number of the step the next sphgpu_advance performs (1 after upload). */
int64_t sphgpu_step(sphgpu_ctx* ctx);

/* Exact checkpoint of a context: every device array (state, integrator
 * history, slot order, cell tables, candidate list, rebuild gate) and the
 * step state, so that a context restored from it continues bit for bit.
 * The reference's nearest counterpart is its particle file
 * (read_initial_file / write_initial_file, input.f90:22-92, through
 * SPH.f90:56-93), which holds rounded fields and no integrator history.
 * The blob is private to this library version; restore refuses another's.
 * sphgpu_checkpoint_size gives the bytes sphgpu_checkpoint_save writes;
 * sphgpu_checkpoint_restore creates a new context (failure message via
 * sphgpu_last_error(NULL), as for sphgpu_create). */
int sphgpu_checkpoint_size(sphgpu_ctx* ctx, int64_t* bytes);
int sphgpu_checkpoint_save(sphgpu_ctx* ctx, void* buf, int64_t bytes);
int sphgpu_checkpoint_restore(const void* buf, int64_t bytes, sphgpu_ctx** out);

/* What a context holds: its config (cfg_out, may be NULL) and counts3 =
 * (n_real, n_dummy, capacity). For a caller that gets a context from
 * sphgpu_checkpoint_restore. No reference counterpart. */
int sphgpu_context_info(sphgpu_ctx* ctx, sphgpu_config* cfg_out, int64_t* counts3);

/* Message of the last failed call on ctx; with ctx == NULL, the message of
 * the last failed sphgpu_create or sphgpu_checkpoint_restore on the calling
 * thread (or of sphgpu_pairs_noalloc). */
const char* sphgpu_last_error(sphgpu_ctx* ctx);

void sphgpu_destroy(sphgpu_ctx* ctx);

/* ------------------------------------------------------------------ batches
 * Many independent runs resident on one GPU, advanced together: every kernel
 * launch covers all runnable runs of the batch through a device run table.
 * No counterpart in the reference, which runs one case per process
 * (src/main.f90:16-22); each run follows the same step as sphgpu_advance
 * (timeIntegration.f90:137-252 with singleStep.f90:58-144), restricted to the
 * fluid set of the fast path (cubic spline, continuity density, Tait/linear
 * EOS, artificial viscosity, gravity, fixed boundary particles).
 *
 * A run's results do not depend on the other runs in the batch: its grid,
 * slot order, candidate-list rebuilds and arithmetic are its own.
 * One batch = one HIP stream = one host thread. */
typedef struct sphgpu_batch sphgpu_batch;

/* Device bytes an arena of these dimensions takes (for sizing it from free
 * HBM). verlet_cap <= 0 selects the default candidate capacity. */
int sphgpu_batch_bytes(int32_t dim, int32_t precision, int64_t slot_capacity,
                       int64_t cell_capacity, int32_t max_runs, int32_t verlet_cap,
                       int64_t* bytes);

/* Allocate an arena of slot_capacity particle slots and cell_capacity cells
 * for up to max_runs resident runs of one dimension and precision. */
int sphgpu_batch_create(int32_t dim, int32_t precision, int64_t slot_capacity,
                        int64_t cell_capacity, int32_t max_runs, int32_t verlet_cap,
                        sphgpu_batch** out);

/* Admit a run (arrays as for sphgpu_upload) that will advance nsteps steps.
 * win (cells per cutoff, 1..4) and skin_frac (candidate skin / cutoff)
 * tune the neighbour search; <= 0 keeps the defaults. Returns
 * SPHGPU_E_CAPACITY, with the batch unchanged, when the arena or the run
 * table has no room; *run receives the run's handle. */
int sphgpu_batch_add(sphgpu_batch* b, const sphgpu_config* cfg, int64_t n_real,
                     int64_t n_dummy, const double* x, const double* v, const double* mass,
                     const double* rho, const double* p, const double* e, const double* c,
                     const double* h, const int32_t* type, const int32_t* state,
                     const int32_t* boundary, int64_t nsteps, int32_t win, double skin_frac,
                     int32_t* run);

/* Advance every runnable run by up to nsteps steps; a run stops at its own
 * last step, and a run that fails (non-finite state, capacity) stops alone
 * with its status recorded. Returns nonzero only for failures of the batch
 * itself. */
int sphgpu_batch_advance(sphgpu_batch* b, int32_t nsteps);

/* State of a run: *next_step = number of the step it would run next;
 * *state = 0 running, 1 finished, or the negative sphgpu_status it failed
 * with (message: sphgpu_batch_run_error). */
int sphgpu_batch_status(sphgpu_batch* b, int32_t run, int64_t* next_step, int32_t* state);
const char* sphgpu_batch_run_error(sphgpu_batch* b, int32_t run);

/* One field of a run, as sphgpu_download. */
int sphgpu_batch_download(sphgpu_batch* b, int32_t run, int32_t field, double* dst);

/* Release a run's slots, cells and table row. */
int sphgpu_batch_remove(sphgpu_batch* b, int32_t run);

/* out8: resident runs, runnable runs, slots in use, slot capacity, cells in
 * use, cell capacity, max runs, candidate capacity per particle. */
int sphgpu_batch_info(sphgpu_batch* b, int64_t* out8);

/* Wall time (HIP events on the batch stream) of the last sphgpu_batch_advance,
 * the steps it ran, and the real-particle steps it advanced over all runs. */
int sphgpu_batch_timing(sphgpu_batch* b, double* ms, int64_t* steps, int64_t* particle_steps);

/* Message of the last failed batch call (b == NULL: of the last failed
 * sphgpu_batch_create on this thread). */
const char* sphgpu_batch_last_error(sphgpu_batch* b);

void sphgpu_batch_destroy(sphgpu_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* SPHGPU_H */
