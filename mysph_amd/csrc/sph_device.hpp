// Device-side data layout and per-particle / per-pair math of the SPH hot
// path for gfx950 (MI355X). Included by sphgpu.hip only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sphgpu.h"

namespace sphgpu {

constexpr double kPi = 3.141592653589793;                  // acos(-1.0_8), tools.f90:8
constexpr double kGmrM1 = 0.2000000476837158203125;        // real(1.2_4,8) - 1, asymmetricSearch.f90:21
constexpr int kNpg = 3;                                      // asymmetricSearch.f90:22
constexpr int kBlock = 256;
constexpr int kWinDefault = 3;   // cells are cutoff/win wide: a +-win cell sweep covers the cutoff
constexpr int kNumXcd = 8;

// error word bits (one word per context, OR-ed by kernels)
constexpr int kErrOutOfRange = 1;
constexpr int kErrNonFinite = 2;
constexpr int kErrPairCap = 4;
constexpr int kErrHRange = 8;    // an evolving smoothing length outgrew the planned search range

// IEEE sum, difference and product, each rounded on its own, for arithmetic
// that must be the reference's (its exact neighbour test, search grid and
// integrator): the compiler may not fuse these into multiply-adds. HIP's
// __dadd_rn and __dmul_rn are plain operators in this build, which hipcc
// contracts into v_fma_f64 like any other a * b + c.
__device__ __forceinline__ double rn_add(double a, double b) {
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ double rn_sub(double a, double b) {
#pragma clang fp contract(off)
    return a - b;
}
__device__ __forceinline__ double rn_mul(double a, double b) {
#pragma clang fp contract(off)
    return a * b;
}

// a + b c with the product and the sum each rounded, as the reference's
// integrator computes it. The leap-frog updates use it in every launch that
// applies them, so a fused correct + predict and the separate launches give
// the same bits.
__device__ __forceinline__ double add_mul(double a, double b, double c) {
#pragma clang fp contract(off)
    return a + b * c;
}
__device__ __forceinline__ float add_mul(float a, float b, float c) {
#pragma clang fp contract(off)
    return a + b * c;
}

struct Material {
    int32_t type, eos;
    double rho0, K, gamma, eta, sigma_y;
};

// Constants of one context, passed by value to every kernel.
struct Params {
    int dim, skf, scale_k;
    int n, n_real;
    int nmat;
    int summation_density, normalize_density, viscosity, gravity, gravity_mode;
    int external_force, arti_visc, arti_stress, kgc, has_solid;
    int has_t100;       // some particle has |type| == 100 (the general force pass keeps one launch)
    int n_sidx;         // real solid targets (general force pass)
    int ext_pow;        // a wall (LJ) or fluid-fluid penalty pair force can fire (their pow terms)
    int mixed;          // FP32 pair physics with density stored as rho - rho0 in FP32
    int pair32;         // general path in mixed precision: FP64 state, FP32 pair arithmetic
    int visc_skip_solid;  // a fluid's viscous pair term skips solid neighbours
    int sle;              // smoothing-length evolution (0 none, 1 from density, 2 from div v)
    int pa_sph;           // pair form of internal_force: 1, 2 or 3 (Riemann)
    int pair_ext;       // external pair forces can act (wall range > 0 or several fluid types)
    double g, av_alpha, av_beta, av_psi;
    double lj_factor, lj_r0, lj_p1, lj_p2;
    double pen_factor, pen_n1, pen_n2, couple_delta;
    double as_lambda, as_delta, as_exponent;
    double dt;
    const Material* mat;   // device copy of the material table (nmat entries)
    // our search grid (fixed per upload; keys are clamped into it)
    double org[3];
    double inv_cs;
    double cs;          // cell size (exactly representable in FP32)
    int nc[3];
    int win;            // sweep half-width in cells
    int64_t ncell;      // nc[0]*nc[1]*nc[2]; cell index ncell holds State == -1 particles
    int nbr_cap;        // neighbour-list slots per particle
    // fluid fast path: candidate (Verlet) list within cutoff + skin, rebuilt
    // when some particle has moved skin/2 since the last build
    double skin;
    double disp_thr2;   // (skin/2)^2, slightly reduced
    double disp_half;   // skin/2, slightly reduced
    int verlet_cap;
    double hmax;        // largest smoothing length (the planned limit when h evolves)
    double hmax_upload; // largest uploaded smoothing length
    // where this run lives: a single-run context has s0 = 0, c0 = 0; a run of
    // a batch owns slots [s0, s0 + n) and cells [c0, c0 + ncell] of the
    // shared arena, and its own grid, gate and error words
    int s0;
    int64_t c0;
    int nblk;                   // blocks of kBlock slots the run spans
    int64_t step0;              // number of the first step of the launch sequence (batch)
    struct RefGrid* grid;
    unsigned long long* vflag;  // the run's rebuild gate (see Gate)
    unsigned long long* vany;   // the batch-wide gate of the shared sort (== vflag for one run)
    int* err;
    int* maxcnt;
    double* partial;            // 6 doubles of bounding-box partials per block of the run
};

// Rebuild gate of the candidate list: flag[0] forces a rebuild, flag[1] is set
// when some particle has moved skin/2 since the last build. A null flag means
// "always".
struct Gate {
    const unsigned long long* flag;
};

__device__ __forceinline__ bool gate_open(const Gate& g) {
    return !g.flag || g.flag[0] != 0ull || g.flag[1] != 0ull;
}

// Reference background grid of one step (asymmetricSearch.f90:49-131).
struct RefGrid {
    double minc[3], maxc[3];
    int num[3];
    int need_win;     // the reference window can exclude a pair (else it need not be checked)
    double pos_err;   // bound on |fl32(x - org) - (x - org)| over all particles
};

// A batch keeps many independent runs in one arena and covers all of them
// with each launch. Per-particle kernels map a launch block to (run row, block
// within the run); slot-strided kernels look the run of a slot up.
struct RunMap {
    const Params* runs;   // device run table (batch launches only)
    const int2* blk;      // launch block -> (run row, block within the run)
    const int* runof;     // slot -> run row; -1 marks a free slot
    const int* rows;      // rows of the active runs (one block each in k_grid)
    int nslot;            // slots below the arena's high-water mark
    int nact;             // particles of the runnable runs
};

template <bool BATCH>
__device__ __forceinline__ const Params& block_run(const Params& P0, const RunMap& M, int b, int& lb) {
    if constexpr (BATCH) {
        const int2 e = M.blk[b];
        lb = e.y;
        return M.runs[e.x];
    } else {
        lb = b;
        return P0;
    }
}

template <bool BATCH>
__device__ __forceinline__ const Params* slot_run(const Params& P0, const RunMap& M, int s) {
    if constexpr (BATCH) {
        const int r = M.runof[s];
        return r < 0 ? nullptr : &M.runs[r];
    } else {
        return &P0;
    }
}

// Set the gate bit of a run (and of the batch) once per wave.
__device__ __forceinline__ void vote_rebuild(const Params& P, bool moved) {
    if (__any(moved) && (threadIdx.x & 63) == 0) {
        if (__atomic_load_n(&P.vflag[1], __ATOMIC_RELAXED) == 0ull) atomicOr(&P.vflag[1], 1ull);
        if (P.vany != P.vflag && __atomic_load_n(&P.vany[1], __ATOMIC_RELAXED) == 0ull)
            atomicOr(&P.vany[1], 1ull);
    }
}

// Neighbour-list layout (candidate lists and exact lists): each group of
// kVlistLanes consecutive slots (one wave) owns a contiguous block of cap rows of kVlistLanes entries; entry k of
// slot s is at row k of its group's block. A wave's load of entry k is one
// contiguous 256-byte row, and its whole list is one contiguous stream, so
// HBM reads it page by page instead of one row per DRAM page.
constexpr int kVlistLanes = 64;

__host__ __device__ __forceinline__ int64_t vlist_at(int s, int k, int cap) {
    return ((int64_t)(s / kVlistLanes) * cap + k) * kVlistLanes + s % kVlistLanes;
}

// entries of a candidate list for `slots` slots with `cap` rows each
inline int64_t vlist_entries(int64_t slots, int cap) {
    return (slots + kVlistLanes - 1) / kVlistLanes * kVlistLanes * cap;
}

__device__ __forceinline__ int xcd_block(int b, int nb) {
    // Work-groups are dispatched round-robin over the XCDs; remap so that each
    // XCD walks a contiguous range of (spatially sorted) particles and its L2
    // holds their neighbours.
    int per = nb / kNumXcd, rem = nb % kNumXcd;
    int x = b % kNumXcd, idx = b / kNumXcd;
    return x * per + min(x, rem) + idx;
}

__device__ __forceinline__ const Material* find_mat(const Params& P, int type) {
    int a = type < 0 ? -type : type;
    for (int k = 0; k < P.nmat; ++k)
        if (P.mat[k].type == a) return &P.mat[k];
    return nullptr;
}

// eos.f90:23-53 (Tait), :149-157 and :219-248 (linear), :57-64 (isothermal).
__device__ __forceinline__ double eos_p(const Material* m, double rho, double p_old) {
    if (!m) return p_old;
    switch (m->eos) {
    case SPHGPU_EOS_TAIT: return m->K * (pow(rho / m->rho0, m->gamma) - 1.0);
    case SPHGPU_EOS_LINEAR: return m->K * (rho - m->rho0);
    case SPHGPU_EOS_ISOTHERMAL: return m->K * rho;
    default: return p_old;
    }
}

// Kernel normalisation factor alpha_d (kernel.f90:47-56, :89-92, :124-133).
// R = double, or float for the pair arithmetic of the mixed general path.
template <int DIM, typename R = double>
__device__ __forceinline__ R kernel_factor(int skf, R h) {
    const R pi = (R)kPi;
    if (skf == 1) {
        if (DIM == 2) return (R)15.0 / ((R)7.0 * pi * (h * h));
        return (R)3.0 / ((R)2.0 * pi * (h * h * h));
    } else if (skf == 2) {
        R sp = sqrt(pi) * h;
        return DIM == 2 ? (R)1.0 / (sp * sp) : (R)1.0 / (sp * sp * sp);
    } else {
        if (DIM == 2) return (R)7.0 / ((R)478.0 * pi * (h * h));
        return (R)1.0 / ((R)120.0 * pi * (h * h * h));
    }
}

// W and dW/dr-over-r: grad W = G * dx (kernel.f90:8-154).
template <int DIM, typename R = double>
__device__ __forceinline__ void kernel_wg(int skf, R r, R h, R factor, R* w, R* G) {
    R q = r / h;
    if (skf == 1) {
        if (q < (R)1.0) {
            *w = factor * ((R)(2.0 / 3.0) - q * q + (q * q * q) / (R)2.0);
            *G = factor * ((R)-2.0 + (R)1.5 * q) / (h * h);
        } else if (q <= (R)2.0) {
            R t = (R)2.0 - q;
            *w = factor * (R)(1.0 / 6.0) * (t * t * t);
            *G = factor * (R)(-0.5) * (t * t) / h / r;
        } else {
            *w = (R)0.0;
            *G = (R)0.0;
        }
    } else if (skf == 2) {
        if (q < (R)3.0) {
            *w = factor * exp(-q * q);
            *G = *w * ((R)-2.0 / (h * h));
        } else {
            *w = (R)0.0;
            *G = (R)0.0;
        }
    } else {
        R a = (R)3.0 - q, b = (R)2.0 - q, c = (R)1.0 - q;
        R a2 = a * a, b2 = b * b, c2 = c * c;
        if (q < (R)1.0) {
            *w = factor * (a2 * a2 * a - (R)6.0 * (b2 * b2 * b) + (R)15.0 * (c2 * c2 * c));
            *G = factor * ((R)-120.0 + (R)120.0 * q - (R)50.0 * (q * q)) / (h * h);
        } else if (q < (R)2.0) {
            *w = factor * (a2 * a2 * a - (R)6.0 * (b2 * b2 * b));
            *G = factor * ((R)-5.0 * (a2 * a2) + (R)30.0 * (b2 * b2)) / h / r;
        } else if (q <= (R)3.0) {
            *w = factor * (a2 * a2 * a);
            *G = factor * ((R)-5.0 * (a2 * a2)) / h / r;
        } else {
            *w = (R)0.0;
            *G = (R)0.0;
        }
    }
}

// Self contribution W(0, h) (density.f90:62).
template <int DIM, typename R = double>
__device__ __forceinline__ R kernel_w0(int skf, R h) {
    R f = kernel_factor<DIM, R>(skf, h);
    if (skf == 1) return f * (R)(2.0 / 3.0);
    if (skf == 2) return f;
    return f * (R)(243.0 - 6.0 * 32.0 + 15.0);
}

// Reference cell of a position, locate (asymmetricSearch.f90:27-47), with
// correctly rounded operations in the reference's order (no contraction).
__device__ __forceinline__ int ref_locate(double x, double mn, double mx, int num, bool* bad) {
    if (!(x <= mx) || !(x >= mn)) {
        *bad = true;
        return 1;
    }
    double t = __ddiv_rn(rn_sub(x, mn), rn_sub(mx, mn));
    t = rn_add(rn_mul(t, (double)num), 1.0);
    return (int)t;
}

__device__ __forceinline__ uint32_t pack_ref(int c0, int c1, int c2) {
    return (uint32_t)(c0 - 1) | ((uint32_t)(c1 - 1) << 10) | ((uint32_t)(c2 - 1) << 20);
}

__device__ __forceinline__ int ref_axis(uint32_t pk, int d) { return (int)((pk >> (10 * d)) & 1023u); }

// Solve A x = b in place of b by the column-pivot elimination of MCPE
// (tools.f90:287-354), including its early return on a singular row.
template <int DIM>
__device__ __forceinline__ void mcpe_solve(const double* Ain, double* b) {
    double A[DIM * DIM], B[DIM], X[DIM];
#pragma unroll
    for (int q = 0; q < DIM * DIM; ++q) A[q] = Ain[q];
#pragma unroll
    for (int q = 0; q < DIM; ++q) { B[q] = b[q]; X[q] = b[q]; }
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        int p = i;
        for (int j = i + 1; j < DIM; ++j)
            if (fabs(A[p * DIM + i]) < fabs(A[j * DIM + i])) p = j;
        for (int j = 0; j < DIM; ++j) {
            double t = A[i * DIM + j]; A[i * DIM + j] = A[p * DIM + j]; A[p * DIM + j] = t;
        }
        double t = B[i]; B[i] = B[p]; B[p] = t;
    }
    for (int i = 0; i < DIM - 1; ++i)
        for (int j = i + 1; j < DIM; ++j) {
            double f = A[j * DIM + i] / A[i * DIM + i];
            double s = 0.0;
            for (int k = i; k < DIM; ++k) {
                A[j * DIM + k] = A[j * DIM + k] - A[i * DIM + k] * f;
                s = s + A[j * DIM + k] * A[j * DIM + k];
            }
            B[j] = B[j] - B[i] * f;
            if (s == 0.0) return;  // b keeps its input value
        }
    X[DIM - 1] = B[DIM - 1] / A[(DIM - 1) * DIM + DIM - 1];
    for (int i = DIM - 2; i >= 0; --i) {
        for (int j = DIM - 1; j >= i + 1; --j) B[i] = B[i] - A[i * DIM + j] * X[j];
        X[i] = B[i] / A[i * DIM + i];
    }
#pragma unroll
    for (int q = 0; q < DIM; ++q) b[q] = X[q];
}

__device__ __forceinline__ double comp(const double4& v, int d) {
    return d == 0 ? v.x : (d == 1 ? v.y : v.z);
}

}  // namespace sphgpu
