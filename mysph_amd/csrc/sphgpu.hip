// libsphgpu: the reference SPH step (NNPS + pair physics + leap-frog update)
// as HIP kernels for gfx950. C ABI in include/sphgpu.h.
//
// Step pipeline on one stream (every array stays resident on the device):
//   k_grid         reference BGGS grid of the step (asymmetricSearch.f90:49-131)
//   k_keys         our cell key (clamped fixed grid) + reference cell per particle,
//                  cell counts
//   k_scan*        exclusive scan of the counts -> cell start table
//   k_scatter      particles into cell order; k_cellsort makes the order
//                  within a cell deterministic (ascending slot)
//   k_point_build  leap-frog half-step predict + EOS (timeIntegration.f90:137-170,
//                  internalForce.f90:103-134) and the neighbour search with the
//                  reference's acceptance rule and window (asymmetricSearch.f90:206-254)
//   k_kgc / k_sumdens / k_strain / k_solid_point   optional passes
//   k_force        continuity, divergence, pressure/viscous/solid stress, wall and
//                  coupling forces, artificial viscosity and stress, gravity
//   k_correct      leap-frog correct (timeIntegration.f90:178-252) + bounding box
//                  partials for the next step's grid
// All sums are gathers (one writer per particle, no floating-point atomics).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "sph_device.hpp"

using namespace sphgpu;

// ----------------------------------------------------------------- device

// Mixed precision: what the pair loop reads of an accepted neighbour, in one
// 32-byte record (one cache line).
struct __align__(32) PairRecF {
    float4 vm;   // vx, vy, vz, m / rho
    float4 th;   // rho, p/rho^2, c, h
};

struct Dev {
    double4* posh;
    double4* vm;
    double* rho;
    double* e;
    double* c;
    double* p;
    int4* meta;
    double4* disp;
    double4* dv;      // dvdt xyz, drhodt w
    double* de;
    double4* vprev;   // v prev xyz, rho prev w
    double* eprev;
    double4* th;      // rho, p/rho^2, c, m/rho
    double2* div;     // divergencePosition, divergenceVelocity
    double* eta;
    double* tds;      // viscous / elastic heating Tds/dt
    double* Ep;       // eta * edot / rho^2 (dim*dim), read by neighbours' viscous term
    double* Ed;       // solid strain rate edot (dim*dim)
    double* rot;      // rotation rate (solid, dim*dim)
    double* shear;    // S
    double* shear_prev;
    double* dsdt;
    double* stress;   // S - p I
    double* Sp;       // stress / rho^2
    double* asR;      // artificial-stress tensor R (2D solids, 4)
    double* kgcA;     // kernel-gradient-correction matrix A (dim*dim)
    double* rho_new;  // summation density result
    double* pold;     // pressure before this step's EOS (SPH = 3 density rate)
    uint32_t* key;
    uint32_t* refcell;
    int* perm;
    int* count;
    int* start;
    int* bsum;
    int* nbr;
    int* ncnt;
    double* partial;
    RefGrid* grid;
    int* err;
    int* maxcnt;
    Material* mat;
    // mixed precision: FP32 copies read by the pair kernels, and FP32 density
    // deviation / derivatives (positions and velocities stay FP64)
    float4* cpos;     // FP32 positions relative to the grid origin, in cell order; w = slot
    // mixed: the candidate test reads only posf, and the k-th candidates of
    // neighbouring lanes are neighbouring slots, so its gathers coalesce best
    // when posf is packed on its own; an accepted neighbour then needs both
    // halves of its 32-byte pair record, which share one cache line
    float4* posf;     // offset from own cell corner (xyz) + packed cell (w)
    PairRecF* prf;    // vm = (vx, vy, vz, m / rho), th = (rho, p/rho^2, c, h)
    float* drho;      // rho - rho0
    float* drho_prev;
    float4* dvf;      // dvdt xyz, d(rho)/dt w
    float* def;       // de/dt
    // fluid fast path
    double4* xb;      // positions at the last candidate-list build
    // mixed precision keeps two output/bookkeeping records in FP32 instead
    float4* dispf;    // displacement (FP64 disp otherwise)
    float4* xbf;      // positions at the last build (FP64 xb otherwise)
    float* ef;        // internal energy (FP64 e otherwise)
    float* eprevf;    // its value at the last correct (FP64 eprev otherwise)
    float2* divf;     // divergences of position and velocity (FP64 div otherwise)
    int* vnbr;        // candidate list, one contiguous block per wave of slots (vlist_at)
    int* vcnt;
    int* ecnt;        // accepted neighbours per particle in the last force pass
    unsigned long long* vflag;   // rebuild gate (see Gate)
    int* sidx;        // general path: the slots of the real solid targets, in slot order
    float* Epf;       // mixed general path (pair32): FP32 copies of Ep and Sp for the pair gathers
    float* Spf;
};

// Block min/max of positions -> partial[blockIdx*6 + (0..2 min, 3..5 max)].
__device__ __forceinline__ void bbox_block_reduce(double mn[3], double mx[3], double* partial,
                                                  int blk) {
    __shared__ double smn[3][kBlock / 64];
    __shared__ double smx[3][kBlock / 64];
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double a = mn[d], b = mx[d];
        for (int o = 32; o > 0; o >>= 1) {
            a = fmin(a, __shfl_xor(a, o));
            b = fmax(b, __shfl_xor(b, o));
        }
        if (lane == 0) { smn[d][wid] = a; smx[d][wid] = b; }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int d = threadIdx.x;
        double a = smn[d][0], b = smx[d][0];
        for (int w = 1; w < kBlock / 64; ++w) { a = fmin(a, smn[d][w]); b = fmax(b, smx[d][w]); }
        partial[blk * 6 + d] = a;
        partial[blk * 6 + 3 + d] = b;
    }
}

// Displacement of the leap-frog correct (timeIntegration.f90:195-212): dt v
// on the first step, accumulated after. FP32 in mixed precision, where it is
// an output-only field.
template <int DIM, bool MIXED>
__device__ __forceinline__ void add_disp(const Dev& D, int s, bool first, double dt, const double4& v) {
    if constexpr (MIXED) {
        float4 ds = D.dispf[s];
        const float ax = (float)(dt * v.x), ay = (float)(dt * v.y), az = DIM == 3 ? (float)(dt * v.z) : 0.0f;
        if (first) {
            ds.x = ax; ds.y = ay; ds.z = az;
        } else {
            ds.x = ds.x + ax; ds.y = ds.y + ay;
            if (DIM == 3) ds.z = ds.z + az;
        }
        D.dispf[s] = ds;
    } else {
        double4 ds = D.disp[s];
        if (first) {
            ds.x = dt * v.x; ds.y = dt * v.y; ds.z = DIM == 3 ? dt * v.z : 0.0;
        } else {
            ds.x = add_mul(ds.x, dt, v.x); ds.y = add_mul(ds.y, dt, v.y);
            if (DIM == 3) ds.z = add_mul(ds.z, dt, v.z);
        }
        D.disp[s] = ds;
    }
}

// Internal energy of the leap-frog correct and the half-step predict that
// follows it, clamped at zero (timeIntegration.f90:147-149, 186-188, 224-226);
// FP32 in mixed precision. Returns nothing: both values are stored.
template <bool MIXED>
__device__ __forceinline__ void energy_correct_predict(const Dev& D, int s, double fac, double h2,
                                                       double dee) {
    if constexpr (MIXED) {
        float e = add_mul(D.eprevf[s], (float)fac, (float)dee);
        if (e < 0) e = 0;
        D.eprevf[s] = e;
        e = add_mul(e, (float)h2, (float)dee);
        if (e < 0) e = 0;
        D.ef[s] = e;
    } else {
        double e = add_mul(D.eprev[s], fac, dee);
        if (e < 0) e = 0;
        D.eprev[s] = e;
        e = add_mul(e, h2, dee);
        if (e < 0) e = 0;
        D.e[s] = e;
    }
}

// Has the particle moved skin/2 since the last candidate-list build? Mixed
// precision keeps the build positions in FP32; its threshold is lowered by
// their rounding error (2^-24 |x| per axis, doubled), so a true move of
// skin/2 always votes. A NaN position votes too.
template <bool MIXED>
__device__ __forceinline__ bool moved_half_skin(const Params& P, const Dev& D, int s, const double4& q) {
    if constexpr (MIXED) {
        const float4 b = D.xbf[s];
        const double ex = q.x - (double)b.x, ey = q.y - (double)b.y, ez = q.z - (double)b.z;
        const double xm = fmax(fabs((double)b.x), fmax(fabs((double)b.y), fabs((double)b.z)));
        const double half = P.disp_half - 2.0 * 1.7320508075688772 * 5.9604644775390625e-8 * xm;
        return !(half > 0.0 && ex * ex + ey * ey + ez * ez < half * half);
    } else {
        const double4 b = D.xb[s];
        const double ex = q.x - b.x, ey = q.y - b.y, ez = q.z - b.z;
        return !(ex * ex + ey * ey + ez * ez < P.disp_thr2);
    }
}

template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_bbox(Params P0, Dev D, RunMap M) {
    int lb;
    const Params& P = block_run<BATCH>(P0, M, blockIdx.x, lb);
    const int l = lb * kBlock + threadIdx.x;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (l < P.n) {
        double4 q = D.posh[P.s0 + l];
        mn[0] = mx[0] = q.x; mn[1] = mx[1] = q.y; mn[2] = mx[2] = q.z;
        if (!isfinite(q.x) || !isfinite(q.y) || !isfinite(q.z)) atomicOr(P.err, kErrNonFinite);
    }
    bbox_block_reduce(mn, mx, P.partial, lb);
}

// Threads of the one block per run of k_grid: its reduction over the block
// partials is a chain of dependent loads per thread, so more threads make it
// shorter.
constexpr int kGridThreads = 1024;

// Reference grid of this step, asymmetricSearch.f90:49-131. One block per run.
template <int DIM, bool BATCH>
__global__ __launch_bounds__(kGridThreads) void k_grid(Params P0, Dev D, RunMap M) {
    const Params& P = BATCH ? M.runs[M.rows[blockIdx.x]] : P0;
    constexpr int kWaves = kGridThreads / 64;
    __shared__ double wmn[3][kWaves], wmx[3][kWaves];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    // the partials of up to kGridUnroll blocks per thread are all loaded
    // before the first min / max (one memory latency rather than a chain); an
    // index past the last block reads the last block again, which changes no
    // min or max
    constexpr int kGridUnroll = 8;
    const double2* __restrict__ part = (const double2*)P.partial;
    for (int b0 = threadIdx.x; b0 < P.nblk; b0 += kGridUnroll * kGridThreads) {
        double2 v[kGridUnroll][3];
#pragma unroll
        for (int u = 0; u < kGridUnroll; ++u) {
            const int b = min(b0 + u * kGridThreads, P.nblk - 1);
#pragma unroll
            for (int q = 0; q < 3; ++q) v[u][q] = part[b * 3 + q];
        }
#pragma unroll
        for (int u = 0; u < kGridUnroll; ++u) {
            const double w[6] = {v[u][0].x, v[u][0].y, v[u][1].x, v[u][1].y, v[u][2].x, v[u][2].y};
            for (int d = 0; d < 3; ++d) {
                mn[d] = fmin(mn[d], w[d]);
                mx[d] = fmax(mx[d], w[3 + d]);
            }
        }
    }
    // min / max are order-independent: wave shuffles, then the block's waves
    for (int d = 0; d < 3; ++d)
        for (int o = 32; o > 0; o >>= 1) {
            mn[d] = fmin(mn[d], __shfl_xor(mn[d], o));
            mx[d] = fmax(mx[d], __shfl_xor(mx[d], o));
        }
    if ((threadIdx.x & 63) == 0)
        for (int d = 0; d < 3; ++d) { wmn[d][threadIdx.x >> 6] = mn[d]; wmx[d][threadIdx.x >> 6] = mx[d]; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double smn[3][1], smx[3][1];
        for (int d = 0; d < 3; ++d) {
            smn[d][0] = wmn[d][0];
            smx[d][0] = wmx[d][0];
            for (int w = 1; w < kWaves; ++w) {
                smn[d][0] = fmin(smn[d][0], wmn[d][w]);
                smx[d][0] = fmax(smx[d][0], wmx[d][w]);
            }
        }
        RefGrid g;
        double delta[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < 3; ++d) { g.minc[d] = 0.0; g.maxc[d] = 0.0; g.num[d] = 1; }
        for (int d = 0; d < DIM; ++d) {
            delta[d] = rn_sub(smx[d][0], smn[d][0]);
            g.maxc[d] = rn_add(smx[d][0], rn_mul(kGmrM1, delta[d]));
            g.minc[d] = rn_sub(smn[d][0], rn_mul(kGmrM1, delta[d]));
        }
        double N = (double)P.n;
        if (DIM == 2) {
            double t = __dsqrt_rn(__ddiv_rn(rn_mul(N, delta[0]), rn_mul(delta[1], (double)kNpg)));
            int c1 = (t >= 0.0 && t < 1e9) ? (int)t + 1 : 0;   // pow(t, 1/2) correctly rounded
            g.num[0] = min(c1, 1000);
            double t2 = __ddiv_rn(rn_mul((double)g.num[0], delta[1]), delta[0]);
            int c2 = (t2 >= 0.0 && t2 < 1e9) ? (int)t2 + 1 : 0;
            g.num[1] = min(c2, 1000);
        } else {
            double t = __ddiv_rn(rn_mul(rn_mul(N, delta[0]), delta[0]),
                                 rn_mul(rn_mul(delta[1], delta[2]), (double)kNpg));
            t = pow(t, 1.0 / 3.0);
            int c1 = (t >= 0.0 && t < 1e9) ? (int)t + 1 : 0;
            g.num[0] = min(c1, 1000);
            double t2 = __ddiv_rn(rn_mul((double)g.num[0], delta[1]), delta[0]);
            double t3 = __ddiv_rn(rn_mul((double)g.num[0], delta[2]), delta[0]);
            int c2 = (t2 >= 0.0 && t2 < 1e9) ? (int)t2 + 1 : 0;
            int c3 = (t3 >= 0.0 && t3 < 1e9) ? (int)t3 + 1 : 0;
            g.num[1] = min(c2, 1000);
            g.num[2] = min(c3, 1000);
        }
        // a flat bounding box divides by zero above (undefined in the
        // reference): report it and keep the grid usable
        for (int d = 0; d < DIM; ++d)
            if (!(delta[d] > 0.0) || g.num[d] < 1) {
                atomicOr(P.err, kErrOutOfRange);
                g.num[d] = 1;
                g.maxc[d] = g.minc[d] + 1.0;
            }
        // window w = int(h a) + 1 cells, reach of the cutoff = scale_k h a cells
        // (a = num/span); it can drop a pair only if scale_k h a > int(h a) + 1,
        // which needs scale_k h a >= 1 for some h <= hmax
        g.need_win = 0;
        for (int d = 0; d < DIM; ++d) {
            double a = g.num[d] / (g.maxc[d] - g.minc[d]);
            if (P.scale_k * P.hmax * a * (1.0 + 1e-9) >= 1.0) g.need_win = 1;
        }
        double X = 0.0;
        for (int d = 0; d < DIM; ++d)
            X = fmax(X, fmax(fabs(smx[d][0] - P.org[d]), fabs(smn[d][0] - P.org[d])));
        g.pos_err = 2.0 * 5.9604644775390625e-8 * X;   // 2^-24 |x - org|, doubled
        *P.grid = g;
    }
}

// Our cell key of slot s (clamped into the fixed grid; the off cell ncell for
// a particle out of the run) counted into its cell, and the reference cell.
template <int DIM>
__device__ __forceinline__ void key_slot(const Params& P, const Dev& D, int s, bool keys_on, bool ref_on) {
    double4 q = D.posh[s];
    int4 mt = D.meta[s];
    uint32_t key;
    if (mt.y == -1) {
        key = (uint32_t)P.ncell;
    } else {
        double xs[3] = {q.x, q.y, q.z};
        int c[3] = {0, 0, 0};
        for (int d = 0; d < DIM; ++d) {
            double t = floor((xs[d] - P.org[d]) * P.inv_cs);
            t = fmin(fmax(t, 0.0), (double)(P.nc[d] - 1));
            c[d] = (int)t;
        }
        key = (uint32_t)((int64_t)c[0] + (int64_t)P.nc[0] * ((int64_t)c[1] + (int64_t)P.nc[1] * c[2]));
    }
    key += (uint32_t)P.c0;
    if (ref_on && mt.y != -1) {
        double xs[3] = {q.x, q.y, q.z};
        const RefGrid g = *P.grid;
        bool bad = false;
        int rc[3] = {1, 1, 1};
        for (int d = 0; d < DIM; ++d) rc[d] = ref_locate(xs[d], g.minc[d], g.maxc[d], g.num[d], &bad);
        if (bad) atomicOr(P.err, (isfinite(q.x) && isfinite(q.y) && isfinite(q.z)) ? kErrOutOfRange
                                                                                  : kErrNonFinite);
        D.refcell[s] = pack_ref(rc[0], rc[1], rc[2]);
    }
    if (keys_on) {
        D.key[s] = key;
        atomicAdd(&D.count[key], 1);
    }
}

template <int DIM, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_keys(Params P0, Dev D, Gate G, int with_ref, RunMap M) {
    // with_ref 1: reference cells with the keys; 2 (one run): reference cells
    // whenever the reference window can exclude pairs, keys only when the
    // gate is open (the fast path's per-step reference cells ride on this
    // launch instead of one of their own)
    const bool keys_on = gate_open(G);
    const bool ref_on = with_ref == 1 || (!BATCH && with_ref == 2 && P0.grid->need_win);
    if (!keys_on && !ref_on) return;
    const int lim = BATCH ? M.nslot : P0.n;
    for (int s = blockIdx.x * kBlock + threadIdx.x; s < lim; s += gridDim.x * kBlock) {
        const Params* Pp = slot_run<BATCH>(P0, M, s);
        if (!Pp || (BATCH && !gate_open(Gate{Pp->vflag}))) continue;   // only runs that rebuild
        key_slot<DIM>(*Pp, D, s, keys_on, ref_on);
    }
}

// Grid of the grid-stride kernels (gated kernels exit at once when no rebuild
// is due, so a small grid keeps that cheap).
constexpr int kStrideBlocks = 2048;

// Exclusive scan of count[0..m) into start[0..m], 2048 entries per block.
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ int block_exclusive_scan(int v, int* total) {
    __shared__ int ws[kBlock / 64];
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) ws[wid] = x;
    __syncthreads();
    int off = 0, tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        if (w < wid) off += ws[w];
        tot += ws[w];
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

__global__ __launch_bounds__(kBlock) void k_scan1(const int* __restrict__ count, int m, int* bsum, Gate G) {
    if (!gate_open(G)) return;
    const int base = blockIdx.x * kScanTile;
    int s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        int i = base + threadIdx.x + k * kBlock;   // coalesced
        if (i < m) s += count[i];
    }
    int tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void k_scan2(int* bsum, int nb, Gate G) {
    if (!gate_open(G)) return;
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += kBlock) {
        int b = b0 + threadIdx.x;
        int v = b < nb ? bsum[b] : 0;
        int tot;
        int ex = block_exclusive_scan(v, &tot);
        if (b < nb) bsum[b] = carry + ex;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[nb] = carry;   // entries sorted in all
}

// Cell start table: start[c] = exclusive prefix of count; the tile goes through
// LDS so that global loads and stores stay coalesced. With own_prefix the
// block adds up the raw tile totals of k_scan1 before it (cheap while there
// are few tiles, and one launch fewer); otherwise bsum was scanned by k_scan2.
// start[base + i] = base_off + exclusive prefix of count within the tile of
// kScanTile entries at base; returns the tile's total.
__device__ __forceinline__ int scan_tile(const int* __restrict__ count, int m, int base, int base_off,
                                         int* start) {
    __shared__ int tile[kScanTile];
    for (int k = 0; k < kScanItems; ++k) {
        int i = threadIdx.x + k * kBlock;
        tile[i] = (base + i < m) ? count[base + i] : 0;
    }
    __syncthreads();
    int v[kScanItems];
    int s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = tile[threadIdx.x * kScanItems + k];
        s += v[k];
    }
    int tot;
    int ex = block_exclusive_scan(s, &tot) + base_off;
    for (int k = 0; k < kScanItems; ++k) {
        tile[threadIdx.x * kScanItems + k] = ex;
        ex += v[k];
    }
    __syncthreads();
    for (int k = 0; k < kScanItems; ++k) {
        int i = threadIdx.x + k * kBlock;
        if (base + i < m) start[base + i] = tile[i];
    }
    __syncthreads();   // the tile is free for the next call
    return tot;
}

__global__ __launch_bounds__(kBlock) void k_scan3(const int* __restrict__ count, int m,
                                                  const int* bsum, int* start, int own_prefix, Gate G) {
    if (!gate_open(G)) return;
    int base_off = 0;
    if (own_prefix) {
        int acc = 0;
        for (int b = threadIdx.x; b < (int)blockIdx.x; b += kBlock) acc += bsum[b];
        block_exclusive_scan(acc, &base_off);
    } else {
        base_off = bsum[blockIdx.x];
    }
    const int tot = scan_tile(count, m, blockIdx.x * kScanTile, base_off, start);
    if (threadIdx.x == 0) {
        if (own_prefix && blockIdx.x == gridDim.x - 1) start[m] = base_off + tot;
        if (!own_prefix && blockIdx.x == 0) start[m] = bsum[gridDim.x];
    }
}

// Tiles up to which k_scan3 adds up its own prefix (its work grows with the
// square of the tile count).
constexpr int kOwnPrefixTiles = 4096;

// Place each particle in its cell. The decrement leaves every count at zero
// again for the next step.
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_scatter(Params P, Dev D, Gate G, RunMap M) {
    if (!gate_open(G)) return;
    const int lim = BATCH ? M.nslot : P.n;
    for (int s = blockIdx.x * kBlock + threadIdx.x; s < lim; s += gridDim.x * kBlock) {
        if (BATCH) {
            const Params* Pp = slot_run<BATCH>(P, M, s);
            if (!Pp || !gate_open(Gate{Pp->vflag})) continue;
        }
        uint32_t key = D.key[s];
        int pos = D.start[key] + atomicSub(&D.count[key], 1) - 1;
        D.perm[pos] = s;
    }
}

// Ascending slot order inside each cell (the pair order, and so every sum, is
// the same from run to run), then the FP32 positions (relative to the run's
// grid origin) in cell order with the slot in w: the candidate loop of the
// search streams these 16-byte records. One thread per cell.
template <bool BATCH>
__device__ __forceinline__ void sort_cell(const Params& P0, const Dev& D, int c, const RunMap& M) {
    const int a = D.start[c], b = D.start[c + 1];
    if (a == b) return;
    int* perm = D.perm;
    for (int i = a + 1; i < b; ++i) {
        int v = perm[i];
        int j = i - 1;
        while (j >= a && perm[j] > v) {
            perm[j + 1] = perm[j];
            --j;
        }
        perm[j + 1] = v;
    }
    const Params& P = BATCH ? M.runs[M.runof[perm[a]]] : P0;   // one run per cell
    for (int k = a; k < b; ++k) {
        const int j = perm[k];
        const double4 q = D.posh[j];
        D.cpos[k] = make_float4((float)(q.x - P.org[0]), (float)(q.y - P.org[1]),
                                (float)(q.z - P.org[2]), __int_as_float(j));
    }
}

template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_cellsort(Params P0, Dev D, int m, Gate G, RunMap M) {
    if (!gate_open(G)) return;
    for (int c = blockIdx.x * kBlock + threadIdx.x; c < m; c += gridDim.x * kBlock) sort_cell<BATCH>(P0, D, c, M);
}

// Runs of at most this many slots and cells (cell table entries, the off cell
// included) sort in one launch of one block, k_sort_small.
constexpr int kSmallSortSlots = 1 << 16;
constexpr int kSmallSortCells = 1 << 18;

// The whole cell sort of a small run in one launch: the reference cells (as
// k_keys) over the launch's blocks, then block 0 alone does the keys, the
// scans of k_scan1/k_scan3, k_scatter and k_cellsort with its barriers
// between them. The same tables as the five launches, in one launch that
// returns at once when no rebuild is due and no reference cells are asked
// for. Cell order within a cell comes from the per-cell sort, so the tables
// do not depend on the order of the counting atomics. Each phase of block 0
// reads what its other threads wrote in the one before: a device-scope fence
// (the writes and atomics are done, the L1 is dropped) and the block barrier
// separate them.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_sort_small(Params P, Dev D, Gate G, int with_ref, int m) {
    const bool keys_on = gate_open(G);
    const bool ref_on = with_ref == 1 || (with_ref == 2 && P.grid->need_win);
    if (!keys_on && !ref_on) return;
    if (ref_on)
        for (int s = blockIdx.x * kBlock + threadIdx.x; s < P.n; s += gridDim.x * kBlock)
            key_slot<DIM>(P, D, s, false, true);
    if (!keys_on || blockIdx.x != 0) return;
    for (int s = threadIdx.x; s < P.n; s += kBlock) key_slot<DIM>(P, D, s, true, false);
    __threadfence();
    __syncthreads();
    int carry = 0;
    for (int base = 0; base < m; base += kScanTile) carry += scan_tile(D.count, m, base, carry, D.start);
    if (threadIdx.x == 0) D.start[m] = carry;
    __threadfence();
    __syncthreads();
    for (int s = threadIdx.x; s < P.n; s += kBlock) {
        const uint32_t key = D.key[s];
        const int pos = D.start[key] + atomicSub(&D.count[key], 1) - 1;
        D.perm[pos] = s;
    }
    __threadfence();
    __syncthreads();
    for (int c = threadIdx.x; c < m; c += kBlock) sort_cell<false>(P, D, c, RunMap{});
}

template <int DIM>
__device__ __forceinline__ void predict_one(const Params& P, const Dev& D, int s, int4 mt,
                                            bool first) {
    double4 v = D.vm[s];
    double rho = D.rho[s], e = D.e[s];
    D.vprev[s] = make_double4(v.x, v.y, v.z, rho);
    D.eprev[s] = e;
    if (first) return;
    double h2 = P.dt / 2;
    double4 dv = D.dv[s];
    e = add_mul(e, h2, D.de[s]);
    if (e < 0) e = 0;
    if (!P.summation_density) rho = add_mul(rho, h2, dv.w);
    v.x = add_mul(v.x, h2, dv.x);
    v.y = add_mul(v.y, h2, dv.y);
    if (DIM == 3) v.z = add_mul(v.z, h2, dv.z);
    D.vm[s] = v;
    D.rho[s] = rho;
    D.e[s] = e;
    if (P.has_solid && mt.x > 100) {
        const Material* m = find_mat(P, mt.x);
        if (!m) return;
        double* S = D.shear + (int64_t)s * DIM * DIM;
        double* Sp = D.shear_prev + (int64_t)s * DIM * DIM;
        const double* dS = D.dsdt + (int64_t)s * DIM * DIM;
        double J2 = 0.0;
        for (int q = 0; q < DIM * DIM; ++q) {
            Sp[q] = S[q];
            S[q] = add_mul(S[q], h2, dS[q]);
        }
        if (m->sigma_y > 0) {
            for (int q = 0; q < DIM * DIM; ++q) J2 = add_mul(J2, S[q], S[q]);
            double sc = fmin(1.0, sqrt(((m->sigma_y * m->sigma_y) / 3) / J2));
            for (int q = 0; q < DIM * DIM; ++q) S[q] = S[q] * sc;
        }
    }
}

// EOS and the per-particle pair inputs th = (rho, p/rho^2, c, m/rho).
__device__ __forceinline__ void eos_th(const Params& P, const Dev& D, int s, int4 mt) {
    double rho = D.rho[s];
    double p = D.p[s];
    if (!(mt.x > 100 || mt.x < -100)) p = eos_p(find_mat(P, mt.x), rho, p);
    D.p[s] = p;
    double m = D.vm[s].w;
    D.th[s] = make_double4(rho, p / (rho * rho), D.c[s], m / rho);
}

// Mixed precision: predict with FP64 velocity/position accumulation and the
// FP32 density deviation drho = rho - rho0.
template <int DIM>
__device__ __forceinline__ void predict_mixed(const Params& P, const Dev& D, int s, bool first) {
    double4 v = D.vm[s];
    float dr = D.drho[s];
    float e = D.ef[s];
    D.vprev[s] = make_double4(v.x, v.y, v.z, v.w);   // mixed: w holds the mass
    D.drho_prev[s] = dr;
    D.eprevf[s] = e;
    if (first) return;
    const double h2 = P.dt / 2;
    float4 dv = D.dvf[s];
    e = add_mul(e, (float)h2, D.def[s]);
    if (e < 0) e = 0;
    dr = add_mul(dr, (float)h2, dv.w);
    v.x = add_mul(v.x, h2, (double)dv.x);
    v.y = add_mul(v.y, h2, (double)dv.y);
    if (DIM == 3) v.z = add_mul(v.z, h2, (double)dv.z);
    D.vm[s] = v;
    D.drho[s] = dr;
    D.ef[s] = e;
}

__device__ __forceinline__ float4 cell_offset_record(const Params& P, const double4& q);

// FP32 EOS on the density deviation: the Tait term (rho/rho0)^gamma - 1 is
// evaluated as a polynomial in u = drho/rho0 (gamma = 7) or expm1(gamma
// log1p(u)), so it does not cancel.
__device__ __forceinline__ void eos_mixed_vals(const Params& P, const Dev& D, int s, int type, float dr,
                                               const double4& v, const double4& q) {
    const Material* m = find_mat(P, type);
    float rho0 = m ? (float)m->rho0 : 0.0f;
    float rho = rho0 + dr;
    float p;
    if (m && m->eos == SPHGPU_EOS_TAIT) {
        float u = dr / rho0;
        float B = (float)m->K;
        if (m->gamma == 7.0)
            p = B * u * (7.0f + u * (21.0f + u * (35.0f + u * (35.0f + u * (21.0f + u * (7.0f + u))))));
        else
            p = B * expm1f((float)m->gamma * log1pf(u));
    } else if (m && m->eos == SPHGPU_EOS_LINEAR) {
        p = (float)m->K * dr;
    } else if (m && m->eos == SPHGPU_EOS_ISOTHERMAL) {
        p = (float)m->K * rho;
    } else {
        p = (float)D.p[s];
    }
    D.p[s] = (double)p;
    float mass = (float)v.w;
    float irho = 1.0f / rho;
    // vm.w = m / rho: the pair loop needs the volume of every neighbour and
    // recovers the mass as V rho (one multiply instead of a reciprocal per pair)
    PairRecF rec;
    rec.vm = make_float4((float)v.x, (float)v.y, (float)v.z, mass * irho);
    rec.th = make_float4(rho, p * irho * irho, (float)D.c[s], (float)q.w);
    D.prf[s] = rec;
    D.posf[s] = cell_offset_record(P, q);
}

// FP32 position record: the offset from the particle's own cell corner, with
// the exact cell size cs: dx of a pair is (off_i - off_j) + (c_i - c_j) cs,
// accurate to ~1e-7 of a cell. w: the packed cell, bit 31 set when the
// particle is outside the grid (its offset is then not bounded by a cell).
__device__ __forceinline__ float4 cell_offset_record(const Params& P, const double4& q) {
    double xs[3] = {q.x, q.y, q.z};
    int c[3] = {0, 0, 0};
    float off[3] = {0.0f, 0.0f, 0.0f};
    bool escaped = false;
    for (int d = 0; d < 3; ++d) {
        if (d < P.dim) {
            double t0 = floor((xs[d] - P.org[d]) * P.inv_cs);
            double t = fmin(fmax(t0, 0.0), (double)(P.nc[d] - 1));
            escaped |= t != t0;
            c[d] = (int)t;
            off[d] = (float)(xs[d] - P.org[d] - t * P.cs);
        }
    }
    // bit 31: the particle is outside the grid (offset not bounded by a cell)
    uint32_t pk = P.dim == 3 ? ((uint32_t)c[0] | ((uint32_t)c[1] << 11) | ((uint32_t)c[2] << 22))
                             : ((uint32_t)c[0] | ((uint32_t)c[1] << 16));
    if (escaped) pk |= 0x80000000u;
    return make_float4(off[0], off[1], off[2], __uint_as_float(pk));
}

template <int DIM>
__device__ __forceinline__ void unpack_cell(uint32_t pk, int& x, int& y, int& z) {
    if (DIM == 3) { x = (int)(pk & 2047u); y = (int)((pk >> 11) & 2047u); z = (int)((pk >> 22) & 511u); }
    else { x = (int)(pk & 65535u); y = (int)((pk >> 16) & 32767u); z = 0; }
}

// x_i - x_j in FP32 from two cell-offset records: (off_i - off_j) + (c_i -
// c_j) cs, the cell part exact.
template <int DIM>
__device__ __forceinline__ void rec_dx(const float4& pi, const float4& pj, float cs, float dx[3]) {
    int cxi, cyi, czi, cxj, cyj, czj;
    unpack_cell<DIM>(__float_as_uint(pi.w), cxi, cyi, czi);
    unpack_cell<DIM>(__float_as_uint(pj.w), cxj, cyj, czj);
    dx[0] = (pi.x - pj.x) + (float)(cxi - cxj) * cs;
    dx[1] = (pi.y - pj.y) + (float)(cyi - cyj) * cs;
    dx[2] = DIM == 3 ? (pi.z - pj.z) + (float)(czi - czj) * cs : 0.0f;
}

// FP32 pair record of the general path in mixed precision (pair32): velocity
// and volume m / rho, then rho, p / rho^2, c and h, from the FP64 state.
__device__ __forceinline__ void write_rec32(const Dev& D, int s) {
    const double4 v = D.vm[s];
    const double4 t = D.th[s];
    PairRecF rec;
    rec.vm = make_float4((float)v.x, (float)v.y, (float)v.z, (float)t.w);
    rec.th = make_float4((float)t.x, (float)t.y, (float)t.z, (float)D.posh[s].w);
    D.prf[s] = rec;
}

__device__ __forceinline__ void eos_mixed(const Params& P, const Dev& D, int s, int type) {
    eos_mixed_vals(P, D, s, type, D.drho[s], D.vm[s], D.posh[s]);
}

// Predict + EOS for every particle, then the neighbour search for real
// particles: candidates from the +-win cell sweep of our grid, accepted with
// the reference rule r < scale_k * h_i (r from the sum of squares in axis order,
// correctly rounded, no contraction), j != i, and j inside i's reference window
// (asymmetricSearch.f90:209-237).
template <int DIM, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_point_build(Params P, Dev D, int mode) {
    // mode bit 0: first step (no predict); bit 1: skip predict entirely
    // (derivatives only); bit 2: count only (no list writes)
    int s = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (s >= P.n) return;
    int4 mt = D.meta[s];
    bool real = mt.w < P.n_real;
    if (MIXED) {
        if (real && !(mode & 2)) predict_mixed<DIM>(P, D, s, mode & 1);
        eos_mixed(P, D, s, mt.x);
    } else {
        if (real && !(mode & 2)) predict_one<DIM>(P, D, s, mt, mode & 1);
        // the Riemann density rate reads the pressures of the last EOS
        // (con_density_riemann runs before internal_force, singleStep.f90:90-120);
        // a count-only sizing pass leaves the state alone, so that the
        // uploaded pressures are the ones the first pass saves
        if (!(mode & 4)) {
            if (P.pa_sph == 3) D.pold[s] = D.p[s];
            if (!P.summation_density) eos_th(P, D, s, mt);
        }
        if (P.pair32) {
            D.posf[s] = cell_offset_record(P, D.posh[s]);
            write_rec32(D, s);
        }
    }
    if (!real || mt.y == -1) {
        if (real) D.ncnt[s] = 0;
        return;
    }
    const double4 qi = D.posh[s];
    const double hi = qi.w;
    const double cut = hi * (double)P.scale_k;
    const uint32_t key = D.key[s];
    const int cx = (int)(key % (uint32_t)P.nc[0]);
    const int rest = (int)(key / (uint32_t)P.nc[0]);
    const int cy = rest % P.nc[1];
    const int cz = rest / P.nc[1];
    const RefGrid g = *D.grid;
    const uint32_t rci = D.refcell[s];
    int wr[3] = {0, 0, 0}, rc[3] = {0, 0, 0};
    bool need_win = false;
    for (int d = 0; d < DIM; ++d) {
        // cellNumPerHSML, asymmetricSearch.f90:209-211
        const double span = g.maxc[d] - g.minc[d];
        wr[d] = (int)(hi / span * g.num[d]) + 1;
        rc[d] = ref_axis(rci, d);
        // the window can only drop a pair if it is narrower than the cutoff
        // measured in reference cells (rounded up, with slack for rounding)
        const double reach = cut / (span / g.num[d]) * (1.0 + 1e-12) + 1e-9;
        if (wr[d] < (int)ceil(reach)) need_win = true;
    }
    // FP32 pre-test: accept if r2 < lo, reject if r2 >= hi, decide the band
    // between in FP64 with the reference's arithmetic. M bounds |r2_f32 - r2|.
    const double u = 5.9604644775390625e-8;
    const double e = g.pos_err + u * cut;
    const double M = 2.0 * 1.7320508075688772 * cut * e + 3.0 * e * e + 3.0 * u * cut * cut;
    const double cut2 = cut * cut;
    const double margin = 2.0 * M + 1e-6 * cut2;
    const float lo = (float)(cut2 - margin), hi2 = (float)(cut2 + margin);
    const float xf = (float)(qi.x - P.org[0]), yf = (float)(qi.y - P.org[1]),
                zf = (float)(qi.z - P.org[2]);
    int cnt = 0;
    const int cap = (mode & 4) ? 0 : P.nbr_cap;
    const int W = P.win;
    const int zlo = DIM == 3 ? max(cz - W, 0) : 0, zhi = DIM == 3 ? min(cz + W, P.nc[2] - 1) : 0;
    const int ylo = max(cy - W, 0), yhi = min(cy + W, P.nc[1] - 1);
    const int xlo = max(cx - W, 0), xhi = min(cx + W, P.nc[0] - 1);
    auto test = [&](const float4 cj) {
        const float dx = xf - cj.x, dy = yf - cj.y, dz = DIM == 3 ? zf - cj.z : 0.0f;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 < hi2)) return;
        const int j = __float_as_int(cj.w);
        if (j == s) return;
        if (!(r2 < lo)) {
            // rounding band: the reference's test, asymmetricSearch.f90:227-237
            const double4 qj = D.posh[j];
            const double dx0 = rn_sub(qi.x, qj.x);
            const double dx1 = rn_sub(qi.y, qj.y);
            double dr = rn_add(rn_mul(dx0, dx0), rn_mul(dx1, dx1));
            if (DIM == 3) {
                const double dx2 = rn_sub(qi.z, qj.z);
                dr = rn_add(dr, rn_mul(dx2, dx2));
            }
            if (!(__dsqrt_rn(dr) < cut)) return;
        }
        if (need_win) {
            const uint32_t rcj = D.refcell[j];
            bool inwin = true;
            for (int d = 0; d < DIM; ++d) inwin &= abs(ref_axis(rcj, d) - rc[d]) <= wr[d];
            if (!inwin) return;
        }
        if (cnt < cap) D.nbr[vlist_at(s, cnt, cap)] = j;
        ++cnt;
    };
    // rows of cells along x are contiguous in cell order: one range per (y, z);
    // the next row's range is fetched while the current one is scanned
    int z = zlo, y = ylo;
    int64_t row = (int64_t)P.nc[0] * ((int64_t)y + (int64_t)P.nc[1] * z);
    int kb = D.start[row + xlo], ke = D.start[row + xhi + 1];
    for (;;) {
        int ny = y + 1, nz = z;
        if (ny > yhi) { ny = ylo; ++nz; }
        const bool more = nz <= zhi;
        int nkb = 0, nke = 0;
        if (more) {
            const int64_t nrow = (int64_t)P.nc[0] * ((int64_t)ny + (int64_t)P.nc[1] * nz);
            nkb = D.start[nrow + xlo];
            nke = D.start[nrow + xhi + 1];
        }
        int k = kb;
        for (; k + 1 < ke; k += 2) {
            const float4 c0 = D.cpos[k];
            const float4 c1 = D.cpos[k + 1];
            test(c0);
            test(c1);
        }
        if (k < ke) test(D.cpos[k]);
        if (!more) break;
        y = ny; z = nz; kb = nkb; ke = nke;
    }
    D.ncnt[s] = cnt;
    if (cnt > P.nbr_cap && !(mode & 4)) atomicOr(D.err, kErrPairCap);
    // same-address atomics serialise: only a count above the stored maximum goes
    if (cnt > __atomic_load_n(D.maxcnt, __ATOMIC_RELAXED)) atomicMax(D.maxcnt, cnt);
}

// Lanes per target of the general pair kernels: 1 (one lane walks a target's
// whole neighbour list), or kLanesPerTarget for runs too small to fill the
// chip with one lane per target and for the few solid targets of a fluid run,
// whose lanes would otherwise walk long lists one dependent gather after
// another. The lanes of a target take every kLanesPerTarget-th neighbour,
// and a fixed shuffle tree adds their partial sums (the same order in every
// run).
constexpr int kLanesPerTarget = 8;
// real particles up to which a run takes kLanesPerTarget lanes per target
constexpr int kSmallRun = 1 << 18;

// Slot of this lane's target (P.n: none) and the lane's place in its group.
// LPT == 1: one lane per slot, XCD-aware block order. LPT > 1: groups of LPT
// lanes over the listed slots (list) or over all slots; a group is whole, so
// every lane of a group takes the same early exits.
template <int LPT>
__device__ __forceinline__ int target_slot(const Params& P, const int* list, int count, int& sub) {
    if (LPT == 1) {
        sub = 0;
        return xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    }
    const int g = blockIdx.x * kBlock + threadIdx.x;
    sub = g % LPT;
    const int t = g / LPT;
    if (list) return t < count ? list[t] : P.n;
    return t;
}

// Sum of v over the LPT lanes of a group, in a fixed tree order.
template <int LPT, typename T>
__device__ __forceinline__ T group_sum(T v) {
    for (int o = LPT / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Pair geometry of target i and neighbour j: dx = x_i - x_j, r, W, grad W.
// R = double, or float in the mixed general path (dx formed in FP64 and
// rounded once; the corrected-gradient solve stays FP64).
template <int DIM, typename R = double>
struct PairGeo {
    R dx[3];
    R r2, r, w, gw[3];
};

// r, W and grad W of a pair whose g.dx is set (hi: the target's h).
template <int DIM, typename R>
__device__ __forceinline__ void pair_geo_dx(const Params& P, R hi, R factor, const double* A,
                                            PairGeo<DIM, R>& g) {
    R dr = g.dx[0] * g.dx[0] + g.dx[1] * g.dx[1];
    if (DIM == 3) dr += g.dx[2] * g.dx[2];
    g.r2 = dr;
    g.r = sqrt(dr);
    R G;
    kernel_wg<DIM, R>(P.skf, g.r, hi, factor, &g.w, &G);
    for (int d = 0; d < 3; ++d) g.gw[d] = (d < DIM) ? G * g.dx[d] : (R)0.0;
    if (A) {
        double b[3] = {(double)g.gw[0], (double)g.gw[1], (double)g.gw[2]};
        mcpe_solve<DIM>(A, b);
        for (int d = 0; d < 3; ++d) g.gw[d] = (R)b[d];
    }
}

template <int DIM, typename R>
__device__ __forceinline__ void pair_geo(const Params& P, const double4& qi, const double4& qj,
                                         R factor, const double* A, PairGeo<DIM, R>& g) {
    g.dx[0] = (R)(qi.x - qj.x);
    g.dx[1] = (R)(qi.y - qj.y);
    g.dx[2] = DIM == 3 ? (R)(qi.z - qj.z) : (R)0.0;
    pair_geo_dx<DIM, R>(P, (R)qi.w, factor, A, g);
}

// kernelGradientCorrection.f90:20-35: A_i = sum_j V_j (x_j - x_i)(e) dW(d).
template <int DIM, typename R, int LPT>
__global__ __launch_bounds__(kBlock) void k_kgc(Params P, Dev D) {
    int sub;
    const int s = target_slot<LPT>(P, nullptr, 0, sub);
    if (s >= P.n) return;
    int4 mt = D.meta[s];
    if (mt.w >= P.n_real || mt.y == -1) return;
    R A[DIM * DIM];
    for (int q = 0; q < DIM * DIM; ++q) A[q] = (R)0.0;
    if (mt.x > 100) {
        double4 qi = D.posh[s];
        R f = kernel_factor<DIM, R>(P.skf, (R)qi.w);
        int n = min(D.ncnt[s], P.nbr_cap);
        int ai = mt.x < 0 ? -mt.x : mt.x;
        for (int k = sub; k < n; k += LPT) {
            int j = D.nbr[vlist_at(s, k, P.nbr_cap)];
            int tj = D.meta[j].x;
            int aj = tj < 0 ? -tj : tj;
            if ((ai - 100) * (aj - 100) < 0) continue;
            double4 qj = D.posh[j];
            PairGeo<DIM, R> g;
            pair_geo<DIM, R>(P, qi, qj, f, nullptr, g);
            R V = (R)(D.vm[j].w / D.rho[j]);
            for (int d = 0; d < DIM; ++d)
                for (int e = 0; e < DIM; ++e) A[d * DIM + e] += (-g.dx[e]) * g.gw[d] * V;
        }
    }
    if (LPT > 1) {
        for (int q = 0; q < DIM * DIM; ++q) A[q] = group_sum<LPT>(A[q]);
        if (sub) return;
    }
    for (int q = 0; q < DIM * DIM; ++q) D.kgcA[(int64_t)s * DIM * DIM + q] = (double)A[q];
}

// density.f90:10-80 (summation density, optional Shepard normalisation).
template <int DIM, typename R, int LPT>
__global__ __launch_bounds__(kBlock) void k_sumdens(Params P, Dev D) {
    int sub;
    const int s = target_slot<LPT>(P, nullptr, 0, sub);
    if (s >= P.n) return;
    int4 mt = D.meta[s];
    if (mt.w >= P.n_real) return;
    double4 qi = D.posh[s];
    R f = kernel_factor<DIM, R>(P.skf, (R)qi.w);
    R w0 = kernel_w0<DIM, R>(P.skf, (R)qi.w);
    R mi = (R)D.vm[s].w;
    // the self terms go to the first lane of a target's group
    R rho = sub == 0 ? mi * w0 : (R)0.0;
    R wi = (R)0.0;
    bool norm = P.normalize_density && D.div[s].x < 1.5;
    if (norm && sub == 0) wi = mi / (R)D.rho[s] * w0;
    int n = (mt.y == -1) ? 0 : min(D.ncnt[s], P.nbr_cap);
    for (int k = sub; k < n; k += LPT) {
        int j = D.nbr[vlist_at(s, k, P.nbr_cap)];
        double4 qj = D.posh[j];
        PairGeo<DIM, R> g;
        pair_geo<DIM, R>(P, qi, qj, f, nullptr, g);
        R mj = (R)D.vm[j].w;
        rho += mj * g.w;
        if (norm) wi += mj / (R)D.rho[j] * g.w;
    }
    if (LPT > 1) {
        rho = group_sum<LPT>(rho);
        wi = group_sum<LPT>(wi);
        if (sub) return;
    }
    if (norm) rho = rho / wi;
    D.rho_new[s] = (double)rho;
}

__global__ __launch_bounds__(kBlock) void k_apply_sumdens(Params P, Dev D) {
    int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= P.n) return;
    int4 mt = D.meta[s];
    if (mt.w < P.n_real) D.rho[s] = D.rho_new[s];
    eos_th(P, D, s, mt);
    if (P.pair32) write_rec32(D, s);
}

// Strain-rate sum of one pair (target s with FP64 state qi, vi; its FP32
// record pfi in mixed precision, R = float, which gathers the neighbour's
// FP32 records; pfj the neighbour's, when the caller has it).
template <int DIM, typename R>
__device__ __forceinline__ void strain_pair(const Params& P, const Dev& D, int j, const float4* pfj_in,
                                            const double4& qi, const double4& vi, const float4& pfi, R f,
                                            const double* A, R T[DIM * DIM]) {
    constexpr bool REC32 = std::is_same<R, float>::value;
    PairGeo<DIM, R> g;
    R V, dvn[3];
    if constexpr (REC32) {
        const float4 pfj = pfj_in ? *pfj_in : D.posf[j];
        const float4 vj = D.prf[j].vm;
        rec_dx<DIM>(pfi, pfj, (float)P.cs, g.dx);
        pair_geo_dx<DIM, R>(P, (R)qi.w, f, A, g);
        V = vj.w;
        dvn[0] = vj.x - (R)vi.x;
        dvn[1] = vj.y - (R)vi.y;
        dvn[2] = vj.z - (R)vi.z;
    } else {
        double4 qj = D.posh[j];
        double4 vj = D.vm[j];
        pair_geo<DIM, R>(P, qi, qj, f, A, g);
        V = (R)D.th[j].w;   // m / rho of this step's EOS
        dvn[0] = (R)(vj.x - vi.x);
        dvn[1] = (R)(vj.y - vi.y);
        dvn[2] = (R)(vj.z - vi.z);
    }
    for (int d = 0; d < DIM; ++d)
        for (int e = 0; e < DIM; ++e) T[d * DIM + e] += V * dvn[d] * g.gw[e];
}

// Strain-rate results of target s from its sums T: the fluid's viscous
// stress term and heating (internalForce.f90:48-112), or the solid's strain
// and rotation rates; eta of the target's material (viscosity.f90:14-33).
template <int DIM, typename R>
__device__ __forceinline__ void strain_finish(const Params& P, const Dev& D, int s, int4 mt, double eta,
                                              const R T[DIM * DIM]) {
    int ai = mt.x < 0 ? -mt.x : mt.x;
    double rho = D.th[s].x;
    int64_t o = (int64_t)s * DIM * DIM;
    if (ai <= 100) {
        double tr = 0.0;
        for (int d = 0; d < DIM; ++d) tr += (double)T[d * DIM + d];
        double sq = 0.0;
        for (int d = 0; d < DIM; ++d)
            for (int e = 0; e < DIM; ++e) {
                double E = 2.0 * (double)T[d * DIM + e] - (d == e ? (2.0 / 3.0) * tr : 0.0);
                sq += E * E;
                D.Ep[o + d * DIM + e] = P.pa_sph == 2 ? eta * E / (rho * rho) : eta * E;
                if (P.pair32) D.Epf[o + d * DIM + e] = (float)D.Ep[o + d * DIM + e];
            }
        D.tds[s] = 0.5 * eta / rho * sq;
    } else {
        for (int d = 0; d < DIM; ++d)
            for (int e = 0; e < DIM; ++e) {
                double E = 0.5 * ((double)T[d * DIM + e] + (double)T[e * DIM + d]);
                D.Ed[o + d * DIM + e] = E;                                                 // edot
                D.rot[o + d * DIM + e] = 0.5 * ((double)T[d * DIM + e] - (double)T[e * DIM + d]);  // rdot
                D.Ep[o + d * DIM + e] = P.pa_sph == 2 ? eta * E / (rho * rho) : eta * E;
                if (P.pair32) D.Epf[o + d * DIM + e] = (float)D.Ep[o + d * DIM + e];
            }
    }
}

// Strain-rate (and solid rotation-rate) sums over the exact list,
// internalForce.f90:48-100, and the fluid viscous heating tdsdt (:108-112).
// T(d,e) = sum_j V_j (v_j - v_i)_d dW_e.
template <int DIM, typename R, int LPT, bool KGC>
__global__ __launch_bounds__(kBlock) void k_strain(Params P, Dev D) {
    int sub;
    const int s = target_slot<LPT>(P, nullptr, 0, sub);
    if (s >= P.n) return;
    int4 mt = D.meta[s];
    const Material* m = find_mat(P, mt.x);
    double eta = (P.viscosity && m) ? m->eta : 0.0;   // viscosity.f90:14-33
    D.eta[s] = eta;
    R T[DIM * DIM];
    for (int q = 0; q < DIM * DIM; ++q) T[q] = (R)0.0;
    bool real = mt.w < P.n_real && mt.y != -1;
    if (real) {
        const double4 qi = D.posh[s];
        const double4 vi = D.vm[s];
        const R f = kernel_factor<DIM, R>(P.skf, (R)qi.w);
        // KGC: the run corrects kernel gradients (compiled out otherwise)
        const double* A = KGC ? D.kgcA + (int64_t)s * DIM * DIM : nullptr;
        float4 pfi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (std::is_same<R, float>::value) pfi = D.posf[s];
        int n = min(D.ncnt[s], P.nbr_cap);
        // the next index is in flight while a pair is summed
        int jn = sub < n ? D.nbr[vlist_at(s, sub, P.nbr_cap)] : 0;
        for (int k = sub; k < n; k += LPT) {
            const int j = jn;
            if (k + LPT < n) jn = D.nbr[vlist_at(s, k + LPT, P.nbr_cap)];
            strain_pair<DIM, R>(P, D, j, nullptr, qi, vi, pfi, f, A, T);
        }
    }
    if (LPT > 1) {
        for (int q = 0; q < DIM * DIM; ++q) T[q] = group_sum<LPT>(T[q]);
        if (sub) return;
    }
    strain_finish<DIM, R>(P, D, s, mt, eta, T);
}

// Solid point pass, internalForce.f90:137-180 and artificialTerm.f90:42-69.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_solid_point(Params P, Dev D) {
    int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= P.n) return;
    int4 mt = D.meta[s];
    int ai = mt.x < 0 ? -mt.x : mt.x;
    if (ai <= 100) return;
    const Material* m = find_mat(P, mt.x);
    int64_t o = (int64_t)s * DIM * DIM;
    double E[DIM * DIM], R[DIM * DIM], S[DIM * DIM];
    for (int q = 0; q < DIM * DIM; ++q) {
        E[q] = P.viscosity ? D.Ed[o + q] : 0.0;
        R[q] = P.viscosity ? D.rot[o + q] : 0.0;
        S[q] = D.shear[o + q];
    }
    double aux = 0.0;
    for (int d = 0; d < DIM; ++d) aux += E[d * DIM + d];
    aux = aux / 3;
    for (int d = 0; d < DIM; ++d) E[d * DIM + d] -= aux;
    double4 t = D.th[s];
    double rho = t.x;
    double sq = 0.0;
    for (int q = 0; q < DIM * DIM; ++q) sq += S[q] * E[q];
    D.tds[s] = sq / rho;
    double p = eos_p(m, rho, D.p[s]);
    D.p[s] = p;
    double eta = (P.viscosity && m) ? m->eta : 0.0;
    for (int d = 0; d < DIM; ++d)
        for (int dd = 0; dd < DIM; ++dd) {
            double v = 2 * eta * E[d * DIM + dd];
            for (int k = 0; k < DIM; ++k)
                v = v + S[d * DIM + k] * R[dd * DIM + k] + S[dd * DIM + k] * R[d * DIM + k];
            if (mt.w < P.n_real) D.dsdt[o + d * DIM + dd] = v;
        }
    double sig[DIM * DIM];
    for (int d = 0; d < DIM; ++d)
        for (int e = 0; e < DIM; ++e) {
            sig[d * DIM + e] = S[d * DIM + e] - (d == e ? p : 0.0);
            D.stress[o + d * DIM + e] = sig[d * DIM + e];
            D.Sp[o + d * DIM + e] = sig[d * DIM + e] / (rho * rho);
            if (P.pair32) D.Spf[o + d * DIM + e] = (float)(sig[d * DIM + e] / (rho * rho));
        }
    D.th[s] = make_double4(rho, p / (rho * rho), t.z, t.w);
    if (P.pair32) write_rec32(D, s);
    if (DIM == 2 && P.arti_stress && mt.x >= 100) {
        double a = sig[0] - sig[3];
        double theta = (a == 0.0) ? 0.5 * kPi : 0.5 * atan((sig[1] + sig[2]) / a);
        double sn = sin(theta), cs = cos(theta);
        double ps0 = cs * cs * sig[0] + sn * sn * sig[3] + 2 * sn * cs * sig[1];
        double ps1 = sn * sn * sig[0] + cs * cs * sig[3] - 2 * sn * cs * sig[1];
        double T0 = ps0 > 0 ? -P.as_lambda * ps0 / (rho * rho) : 0.0;
        double T1 = ps1 > 0 ? -P.as_lambda * ps1 / (rho * rho) : 0.0;
        double* Rs = D.asR + (int64_t)s * 4;
        Rs[0] = cs * cs * T0 + sn * sn * T1;
        Rs[3] = sn * sn * T0 + cs * cs * T1;
        Rs[1] = sn * cs * (T0 - T1);
        Rs[2] = Rs[1];
    }
}

// Smoothing-length evolution after the force pass (singleStep.f90:136,
// hsml.f90:8-38), real particles: sle 1 h = 2 (m/rho)^(1/dim) with the
// density of this step; sle 2 h += dt (h/dim) divergenceVelocity, undone if
// it would reach h <= 0. h beyond the planned range fails the step.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_hsml(Params P, Dev D) {
    int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= P.n) return;
    if (D.meta[s].w >= P.n_real) return;
    double4 q = D.posh[s];
    double h = q.w;
    if (P.sle == 1) {
        h = 2.0 * pow(D.vm[s].w / D.rho[s], 1.0 / DIM);
    } else if (P.sle == 2) {
        const double dh = (q.w / DIM) * D.div[s].y;
        h = q.w + P.dt * dh;
        if (h <= 0) h = h - P.dt * dh;
    }
    if (!(h <= P.hmax)) atomicOr(D.err, kErrHRange);
    q.w = h;
    D.posh[s] = q;
}

// Target classes of the general force pass: a launch for every target, or one
// for fluid targets (|type| < 100) and one for solid targets (|type| > 100),
// each compiled with only its own terms (fewer live registers, more waves in
// flight). |type| == 100 (a fluid that artificial stress also acts on) keeps
// the launch for every target.
constexpr int kTgtAll = 0, kTgtFluid = 1, kTgtSolid = 2;

// The force pass: everything single_step adds into Delta (singleStep.f90:139-144)
// plus continuity density and the divergences. TGT: target class; SPH: 2
// compiles the SPH = 2 form alone, 0 takes P.pa_sph; LPT: lanes per target
// (the solid class takes the listed solid slots); POW: the wall (LJ) and
// fluid-fluid penalty forces are compiled in (a run where neither can fire,
// P.ext_pow == 0, leaves their pow terms out).
template <int DIM, typename R, int TGT, int SPH, int LPT, bool POW>
__global__ __launch_bounds__(kBlock) void k_force(Params P, Dev D) {
    // mixed precision (R = float) gathers the FP32 pair records; FP64 the
    // FP64 state. Every per-pair value is an R, so the FP64 arithmetic is
    // exactly that of the FP64 records.
    constexpr bool REC32 = std::is_same<R, float>::value;
    int sub;
    const int s = target_slot<LPT>(P, TGT == kTgtSolid ? D.sidx : nullptr, P.n_sidx, sub);
    if (s >= P.n) return;
    int4 mt = D.meta[s];
    if (mt.w >= P.n_real) return;
    const int ti = mt.x;
    const int ai = ti < 0 ? -ti : ti;
    if (TGT == kTgtFluid && ai >= 100) return;
    if (TGT == kTgtSolid && ai <= 100) return;
    const bool fluid = TGT == kTgtFluid ? true : TGT == kTgtSolid ? false : ai <= 100;
    const int pa_sph = SPH ? SPH : P.pa_sph;
    const double4 qi = D.posh[s];
    const double4 vi = D.vm[s];
    const double4 thi = D.th[s];
    const R hi = (R)qi.w;
    const R vi_r[3] = {(R)vi.x, (R)vi.y, (R)vi.z};
    const R rhoi = (R)thi.x, pri = (R)thi.y, ci = (R)thi.z;
    float4 pfi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (REC32) pfi = D.posf[s];
    const float csf = (float)P.cs;
    const R f = kernel_factor<DIM, R>(P.skf, hi);
    // kernel gradient correction has solid targets only (create refuses it
    // with fluid real particles)
    const double* A = (TGT != kTgtFluid && P.kgc) ? D.kgcA + (int64_t)s * DIM * DIM : nullptr;
    R Ei[DIM * DIM], Si[DIM * DIM], Ri[4] = {0, 0, 0, 0};
    const bool visc = P.viscosity && fluid;
    for (int q = 0; q < DIM * DIM; ++q) {
        Ei[q] = visc ? (R)D.Ep[(int64_t)s * DIM * DIM + q] : (R)0.0;
        Si[q] = (!fluid) ? (R)D.Sp[(int64_t)s * DIM * DIM + q] : (R)0.0;
    }
    const bool as_on = TGT != kTgtFluid && DIM == 2 && P.arti_stress && ti >= 100;
    R wdelta = (R)1.0;
    if (as_on) {
        for (int q = 0; q < 4; ++q) Ri[q] = (ti >= 100) ? (R)D.asR[(int64_t)s * 4 + q] : (R)0.0;
        R G;
        kernel_wg<DIM, R>(P.skf, (R)P.as_delta, hi, f, &wdelta, &G);
    }
    R acc[3] = {0.0, 0.0, 0.0};
    R de_in = 0.0, de_av = 0.0, sdiv_v = 0.0, sdiv_x = 0.0, drho_r = 0.0;
    int n = (mt.y == -1) ? 0 : min(D.ncnt[s], P.nbr_cap);
    // the neighbour type is read by the viscous, wall/coupling and artificial
    // stress terms; each pair's loads issue together, and the next index is
    // in flight while a pair is summed
    const bool need_type = visc || P.external_force || as_on;
    const bool need_t = visc || !fluid;
    int jn = sub < n ? D.nbr[vlist_at(s, sub, P.nbr_cap)] : 0;
    for (int k = sub; k < n; k += LPT) {
        const int j = jn;
        if (k + LPT < n) jn = D.nbr[vlist_at(s, k + LPT, P.nbr_cap)];
        const int tj = need_type ? D.meta[j].x : 0;
        // the neighbour's viscous (fluid target) or stress (solid target)
        // tensor, FP32 copies in mixed precision
        R Tj[DIM * DIM];
        if (need_t)
            for (int q = 0; q < DIM * DIM; ++q) {
                if constexpr (REC32) Tj[q] = (fluid ? D.Epf : D.Spf)[(int64_t)j * DIM * DIM + q];
                else Tj[q] = (R)(fluid ? D.Ep : D.Sp)[(int64_t)j * DIM * DIM + q];
            }
        PairGeo<DIM, R> g;
        R vj_r[3], hj, mj, rhoj, prj, cj, Vj;
        if constexpr (REC32) {
            const float4 pfj = D.posf[j];
            const PairRecF rec = D.prf[j];
            rec_dx<DIM>(pfi, pfj, csf, g.dx);
            pair_geo_dx<DIM, R>(P, hi, f, A, g);
            vj_r[0] = rec.vm.x; vj_r[1] = rec.vm.y; vj_r[2] = rec.vm.z;
            Vj = rec.vm.w;
            rhoj = rec.th.x; prj = rec.th.y; cj = rec.th.z; hj = rec.th.w;
            mj = Vj * rhoj;
        } else {
            const double4 qj = D.posh[j];
            const double4 vj = D.vm[j];
            const double4 thj = D.th[j];
            pair_geo<DIM, R>(P, qi, qj, f, A, g);
            vj_r[0] = (R)vj.x; vj_r[1] = (R)vj.y; vj_r[2] = (R)vj.z;
            Vj = (R)thj.w;
            rhoj = (R)thj.x; prj = (R)thj.y; cj = (R)thj.z; hj = (R)qj.w;
            mj = (R)vj.w;
        }
        R dv[3] = {vi_r[0] - vj_r[0], vi_r[1] - vj_r[1], DIM == 3 ? vi_r[2] - vj_r[2] : (R)0.0};
        R dvg = dv[0] * g.gw[0] + dv[1] * g.gw[1] + (DIM == 3 ? dv[2] * g.gw[2] : (R)0.0);
        R dxg = g.dx[0] * g.gw[0] + g.dx[1] * g.gw[1] + (DIM == 3 ? g.dx[2] * g.gw[2] : (R)0.0);
        // continuity (density.f90:98-101) and divergences (divergence.f90:17-27)
        sdiv_v += Vj * dvg;
        sdiv_x += Vj * dxg;
        // internal force, SPH = 1 (internalForce.f90:197-225) and SPH = 3,
        // PVRS Riemann (:277-325): fluids only, solids take none
        if (pa_sph != 2) {
            if (fluid) {
                const R rhoij = (R)1.0 / (rhoi * rhoj);
                const R pi = (R)D.p[s], pj = (R)D.p[j];
                R a13, w[3];
                R ex[3] = {0, 0, 0}, v_l = 0, v_r = 0;
                if (pa_sph == 3) {
                    for (int d = 0; d < DIM; ++d) {
                        ex[d] = -g.dx[d] / g.r;   // (x_j - x_i) / |x_j - x_i|
                        v_l += vi_r[d] * ex[d];
                        v_r += vj_r[d] * ex[d];
                    }
                }
                if (pa_sph == 1) {
                    a13 = (pi + pj) * rhoij;
                    for (int d = 0; d < DIM; ++d) w[d] = dv[d];
                } else {
                    const R Zl = rhoi * ci, Zr = rhoj * cj;
                    const R v_ij = (Zl * v_l + Zr * v_r + (pi - pj)) / (Zl + Zr);
                    const R p_star = (Zl * pj + Zr * pi + Zl * Zr * (v_l - v_r)) / (Zl + Zr);
                    a13 = (R)2 * p_star * rhoij;
                    for (int d = 0; d < DIM; ++d) {
                        const R vs = v_ij * ex[d] + ((vi_r[d] + vj_r[d]) / (R)2 - ((v_l + v_r) / (R)2) * ex[d]);
                        w[d] = (R)2 * (vi_r[d] - vs);
                    }
                }
                const bool vj_on = visc && !(P.visc_skip_solid && (tj > 100 || tj < -100));
                R dot = (R)0;
                for (int d = 0; d < DIM; ++d) {
                    R mv = (R)0.0;   // matmul(eta_i edot_i + eta_j edot_j, dwdx)
                    if (vj_on)
                        for (int e = 0; e < DIM; ++e) mv += (Ei[d * DIM + e] + Tj[d * DIM + e]) * g.gw[e];
                    acc[d] += mj * (-a13 * g.gw[d] + mv * rhoij);
                    dot += w[d] * g.gw[d];
                }
                de_in += mj * a13 * dot;
            }
            if (pa_sph == 3) {
                // density.f90:109-145 with the pressures of the last EOS and
                // the neighbour's sound speed in both impedances
                const R Zl = rhoi * cj, Zr = rhoj * cj;
                const R pi = (R)D.pold[s], pj = (R)D.pold[j];
                R v_l = 0, v_r = 0, ex[3] = {0, 0, 0};
                for (int d = 0; d < DIM; ++d) {
                    ex[d] = -g.dx[d] / g.r;
                    v_l += vi_r[d] * ex[d];
                    v_r += vj_r[d] * ex[d];
                }
                const R v_ij = (Zl * v_l + Zr * v_r + (pi - pj)) / (Zl + Zr);
                R dot = (R)0;
                for (int d = 0; d < DIM; ++d) {
                    const R vs = v_ij * ex[d] + ((vi_r[d] + vj_r[d]) / (R)2 - ((v_l + v_r) / (R)2) * ex[d]);
                    dot += (vi_r[d] - vs) * g.gw[d];
                }
                drho_r += (R)2 * rhoi * mj / rhoj * dot;
            }
        } else {
        // internal force, SPH = 2 (internalForce.f90:228-274)
        R aux = pri + prj;
        if (fluid) {
            const bool vj_on = visc && !(P.visc_skip_solid && (tj > 100 || tj < -100));
            for (int d = 0; d < DIM; ++d) {
                R mv = (R)0.0;
                if (vj_on)
                    for (int e = 0; e < DIM; ++e) mv += (Ei[d * DIM + e] + Tj[d * DIM + e]) * g.gw[e];
                acc[d] += mj * (-aux * g.gw[d] + mv);
            }
        } else {
            for (int d = 0; d < DIM; ++d) {
                R sv = (R)0.0;
                for (int e = 0; e < DIM; ++e) sv += (Si[d * DIM + e] + Tj[d * DIM + e]) * g.gw[e];
                acc[d] += mj * sv;
            }
        }
        de_in += mj * aux * dvg;
        }
        // external pair forces (externalForce.f90:103-174)
        if (P.external_force) {
            if (POW && ti > 0 && tj < 0 && (double)g.r < P.lj_r0) {
                R rr = (R)P.lj_r0 / g.r;
                R fl = (R)P.lj_factor * (pow(rr, (R)P.lj_p1) - pow(rr, (R)P.lj_p2)) / (g.r * g.r);
                for (int d = 0; d < DIM; ++d) acc[d] += fl * g.dx[d];
            }
            if (ti != tj && ti > 0 && tj > 0) {
                if (ti < 100 && tj < 100) {
                    // (without POW the run has one fluid type: never taken)
                    R pe = (hi + hj) / ((R)2 * g.r);
                    if (POW && pe >= (R)1) {
                        R fp = (R)P.pen_factor * (pow(pe, (R)P.pen_n1) - pow(pe, (R)P.pen_n2)) / (g.r * g.r);
                        for (int d = 0; d < DIM; ++d) acc[d] += fp * g.dx[d];
                    }
                } else if (!(ti > 100 && tj > 100)) {
                    R et = g.r / (R)0.75 / ((hi + hj) * (R)0.5);
                    R chi = (R)1 - g.r / (R)P.couple_delta;
                    if (chi < (R)0 || chi > (R)1) chi = (R)0;
                    R fc = (R)0;
                    if (et <= (R)0 || et >= (R)2) fc = (R)0;
                    else if (et <= (R)(2.0 / 3)) fc = (R)(2.0 / 3);
                    else if (et <= (R)1) fc = (R)2 * et - (R)1.5 * et * et;
                    else if (et < (R)2) fc = (R)0.5 * (((R)2 - et) * ((R)2 - et));
                    R fcp = (R)((double)0.01f * (thi.z * thi.z)) * chi * fc / (g.r * g.r);
                    for (int d = 0; d < DIM; ++d) acc[d] += fcp * g.dx[d];
                }
            }
        }
        // Monaghan artificial viscosity (artificialViscosity.f90:80-110)
        if (P.arti_visc) {
            R xv = g.dx[0] * dv[0] + g.dx[1] * dv[1] + (DIM == 3 ? g.dx[2] * dv[2] : (R)0.0);
            if (xv < (R)0.0) {
                R hij = (R)0.5 * (hi + hj);
                R rhoij = (R)0.5 * (rhoi + rhoj);
                R cij = (R)0.5 * (ci + cj);
                R ph = (R)P.av_psi * hij;
                R phi = hij * xv / (g.r2 + ph * ph);
                R PI_ij = ((R)P.av_beta * phi - (R)P.av_alpha * cij) * phi / rhoij;
                for (int d = 0; d < DIM; ++d) acc[d] -= mj * PI_ij * g.gw[d];
                de_av += (R)0.5 * mj * PI_ij * dvg;
            }
        }
        // artificial stress (artificialTerm.f90:72-93)
        if (as_on) {
            if (tj >= 100) {
                R F = pow(g.w / wdelta, (R)P.as_exponent);
                const double* Rj = D.asR + (int64_t)j * 4;
                for (int d = 0; d < 2; ++d) {
                    R sv = (R)0.0;
                    for (int e = 0; e < 2; ++e) sv += (Ri[d * 2 + e] + (R)Rj[d * 2 + e]) * g.gw[e];
                    acc[d] += mj * F * sv;
                }
            }
        }
    }
    if (LPT > 1) {
        for (int d = 0; d < 3; ++d) acc[d] = group_sum<LPT>(acc[d]);
        de_in = group_sum<LPT>(de_in);
        de_av = group_sum<LPT>(de_av);
        sdiv_v = group_sum<LPT>(sdiv_v);
        sdiv_x = group_sum<LPT>(sdiv_x);
        drho_r = group_sum<LPT>(drho_r);
        if (sub) return;
    }
    double ac[3] = {(double)acc[0], (double)acc[1], (double)acc[2]};
    if (P.gravity && !(P.gravity_mode == SPHGPU_GRAVITY_SKIP_SOLID && ti > 100))
        ac[DIM - 1] += P.g;
    if (mt.z == 1) ac[0] = ac[1] = ac[2] = 0.0;   // fixedBoundary, boundaryCondition.f90:32-45
    double drho = P.summation_density ? 0.0 : thi.x * (double)sdiv_v;
    if (pa_sph == 3) drho = (double)drho_r;
    D.dv[s] = make_double4(ac[0], ac[1], ac[2], drho);
    double tds = (P.viscosity || !fluid) ? D.tds[s] : 0.0;
    D.de[s] = tds + 0.5 * (double)de_in + (double)de_av;
    D.div[s] = make_double2(-(double)sdiv_x, -(double)sdiv_v);
}

// Fluid force pass for the dam-break physics set: cubic-spline kernel,
// continuity density, SPH = 2 pressure force, Monaghan viscosity, gravity,
// wall/penalty/coupling pair forces, divergences. R = double (FP64) or float
// (mixed: the pair arithmetic in FP32, dx formed in FP64 then rounded).
template <typename R> struct Vec4;
template <> struct Vec4<double> { using T = double4; };
template <> struct Vec4<float> { using T = float4; };

__device__ __forceinline__ double fast_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double fast_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }

template <int DIM, bool MIXED, bool PAIR_EXT>
__global__ __launch_bounds__(kBlock) void k_force_fluid(Params P, Dev D) {
    using R = typename std::conditional<MIXED, float, double>::type;
    using R4 = typename Vec4<R>::T;
    int s = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (s >= P.n) return;
    const int4 mt = D.meta[s];
    if (mt.w >= P.n_real) return;
    const int ti = mt.x;
    R4 vi, thi;
    R pix, piy, piz, hi;
    int cxi = 0, cyi = 0, czi = 0;
    double4 qi;
    if (MIXED) {
        vi = *(const R4*)&D.prf[s].vm;
        thi = *(const R4*)&D.prf[s].th;
        const float4 pf = D.posf[s];
        pix = pf.x; piy = pf.y; piz = pf.z;
        unpack_cell<DIM>(__float_as_uint(pf.w), cxi, cyi, czi);
        hi = thi.w;
    } else {
        vi = *(const R4*)&D.vm[s];
        thi = *(const R4*)&D.th[s];
        qi = D.posh[s];
        hi = (R)qi.w;
    }
    const R csf = (R)P.cs;
    const R inv_h = (R)1 / hi;
    // cubic spline (kernel.f90:30-71): grad W = G dx with
    //   G = f/h^2 (-2 + 1.5 q)        q < 1
    //   G = -f/(2h) (2 - q)^2 / r     1 <= q <= 2
    const R fac = DIM == 2 ? (R)(15.0 / (7.0 * kPi)) * inv_h * inv_h
                           : (R)(3.0 / (2.0 * kPi)) * inv_h * inv_h * inv_h;
    const R g1 = fac * inv_h * inv_h;
    const R g2 = (R)-0.5 * fac * inv_h;
    const R alpha = (R)P.av_alpha, beta = (R)P.av_beta, psi2 = (R)(P.av_psi * P.av_psi);
    R ax = 0, ay = 0, az = 0, de_in = 0, de_av = 0, sv = 0, sx = 0;
    const int n = (mt.y == -1) ? 0 : min(D.ncnt[s], P.nbr_cap);
    const int* __restrict__ nl = D.nbr + vlist_at(s, 0, P.nbr_cap);
    int jn = n > 0 ? nl[0] : 0;
    for (int k = 0; k < n; ++k) {
        const int j = jn;
        if (k + 1 < n) jn = nl[(int64_t)(k + 1) * kVlistLanes];
        R4 vj, thj;
        R dx, dy, dz, hj, Vj;
        if (MIXED) {
            vj = *(const R4*)&D.prf[j].vm;
            thj = *(const R4*)&D.prf[j].th;
            const float4 pf = D.posf[j];
            int cxj, cyj, czj;
            unpack_cell<DIM>(__float_as_uint(pf.w), cxj, cyj, czj);
            dx = (pix - pf.x) + (R)(cxi - cxj) * csf;
            dy = (piy - pf.y) + (R)(cyi - cyj) * csf;
            dz = DIM == 3 ? (piz - pf.z) + (R)(czi - czj) * csf : (R)0;
            hj = thj.w;
            Vj = vj.w;              // vm.w = m / rho
            vj.w = Vj * thj.x;      // mass
        } else {
            vj = *(const R4*)&D.vm[j];
            thj = *(const R4*)&D.th[j];
            const double4 qj = D.posh[j];
            dx = (R)(qi.x - qj.x);
            dy = (R)(qi.y - qj.y);
            dz = DIM == 3 ? (R)(qi.z - qj.z) : (R)0;
            hj = (R)qj.w;
            Vj = thj.w;
        }
        const R r2 = dx * dx + dy * dy + dz * dz;
        const R r = fast_sqrt(r2);
        const R q = r * inv_h;
        const R t = (R)2 - q;
        const R G = q < (R)1 ? g1 * ((R)-2 + (R)1.5 * q) : g2 * t * t * fast_rcp(r);
        const R gx = G * dx, gy = G * dy, gz = G * dz;
        const R dvx = vi.x - vj.x, dvy = vi.y - vj.y, dvz = DIM == 3 ? vi.z - vj.z : (R)0;
        const R dvg = dvx * gx + dvy * gy + dvz * gz;
        const R dxg = dx * gx + dy * gy + dz * gz;
        const R mj = vj.w;
        sv += Vj * dvg;             // continuity / div v (density.f90:98-101)
        sx += Vj * dxg;             // div x (divergence.f90:17-27)
        const R aux = thi.y + thj.y;  // p_i/rho_i^2 + p_j/rho_j^2 (internalForce.f90:236)
        // Monaghan viscosity when approaching (artificialViscosity.f90:80-110)
        const R xv = dx * dvx + dy * dvy + dz * dvz;
        const R hij = (R)0.5 * (hi + hj);
        const R phi = hij * xv * fast_rcp(r2 + psi2 * hij * hij);
        const R Pi = xv < (R)0 ? (beta * phi - alpha * (R)0.5 * (thi.z + thj.z)) * phi *
                                     fast_rcp((R)0.5 * (thi.x + thj.x))
                               : (R)0;
        const R cf = mj * (aux + Pi);
        ax -= cf * gx;
        ay -= cf * gy;
        az -= cf * gz;
        de_in += mj * aux * dvg;
        de_av += (R)0.5 * mj * Pi * dvg;
        if (PAIR_EXT) {
            const int tj = D.meta[j].x;
            R fe = 0;
            if (ti > 0 && tj < 0 && (double)r < P.lj_r0) {
                R rr = (R)P.lj_r0 / r;
                fe += (R)P.lj_factor * ((R)pow((double)rr, P.lj_p1) - (R)pow((double)rr, P.lj_p2)) / r2;
            }
            if (ti != tj && ti > 0 && tj > 0) {
                if (ti < 100 && tj < 100) {
                    R pe = (hi + hj) / ((R)2 * r);
                    if (pe >= (R)1)
                        fe += (R)P.pen_factor * ((R)pow((double)pe, P.pen_n1) - (R)pow((double)pe, P.pen_n2)) / r2;
                } else if (!(ti > 100 && tj > 100)) {
                    R et = r / (R)0.75 / ((hi + hj) * (R)0.5);
                    R chi = (R)1 - r / (R)P.couple_delta;
                    if (chi < 0 || chi > 1) chi = 0;
                    R fc = 0;
                    if (et <= 0 || et >= 2) fc = 0;
                    else if (et <= (R)(2.0 / 3)) fc = (R)(2.0 / 3);
                    else if (et <= 1) fc = 2 * et - (R)1.5 * et * et;
                    else if (et < 2) fc = (R)0.5 * ((2 - et) * (2 - et));
                    fe += (R)0.01f * (thi.z * thi.z) * chi * fc / r2;
                }
            }
            ax += fe * dx;
            ay += fe * dy;
            az += fe * dz;
        }
    }
    if (P.gravity && !(P.gravity_mode == SPHGPU_GRAVITY_SKIP_SOLID && ti > 100)) {
        if (DIM == 2) ay += (R)P.g;
        else az += (R)P.g;
    }
    if (mt.z == 1) ax = ay = az = 0;
    const R drho = thi.x * sv;
    const R de = (R)0.5 * de_in + de_av;
    if (MIXED) {
        D.dvf[s] = make_float4((float)ax, (float)ay, (float)az, (float)drho);
        D.def[s] = (float)de;
    } else {
        D.dv[s] = make_double4(ax, ay, az, drho);
        D.de[s] = de;
    }
    if (MIXED) D.divf[s] = make_float2(-(float)sx, -(float)sv);
    else D.div[s] = make_double2(-(double)sx, -(double)sv);
}

#include "fluid_fast.hpp"
#include "general_vlist.hpp"

// Leap-frog correct (timeIntegration.f90:178-252) and bounding-box partials
// of the new positions for the next step's grid.
// Leap-frog correct of one real particle (timeIntegration.f90:178-252): the
// velocity, density and energy from their values at the last predict, the
// position, and the deviatoric stress of a solid (J2 cap). q: the particle's
// position record, updated.
template <int DIM, bool MIXED>
__device__ __forceinline__ void correct_one(const Params& P, const Dev& D, int s, int4 mt, bool first,
                                            double4& q) {
    const double dt = P.dt;
    const double fac = first ? dt / 2 : dt;
    double4 vp = D.vprev[s];
    double dvx, dvy, dvz, dee;
    bool finite_rho;
    if (MIXED) {
        float4 dv = D.dvf[s];
        dvx = dv.x; dvy = dv.y; dvz = dv.z; dee = D.def[s];
        float dr = add_mul(D.drho_prev[s], (float)fac, dv.w);
        D.drho[s] = dr;
        finite_rho = isfinite(dr);
    } else {
        double4 dv = D.dv[s];
        dvx = dv.x; dvy = dv.y; dvz = dv.z; dee = D.de[s];
        if (!P.summation_density) D.rho[s] = add_mul(vp.w, fac, dv.w);
        finite_rho = isfinite(D.rho[s]);
    }
    if (MIXED) {
        float e = add_mul(D.eprevf[s], (float)fac, (float)dee);
        if (e < 0) e = 0;
        D.ef[s] = e;
    } else {
        double e = add_mul(D.eprev[s], fac, dee);
        if (e < 0) e = 0;
        D.e[s] = e;
    }
    double4 v = D.vm[s];
    v.x = add_mul(vp.x, fac, dvx);
    v.y = add_mul(vp.y, fac, dvy);
    if (DIM == 3) v.z = add_mul(vp.z, fac, dvz);
    D.vm[s] = v;
    add_disp<DIM, MIXED>(D, s, first, dt, v);
    q.x = add_mul(q.x, dt, v.x);
    q.y = add_mul(q.y, dt, v.y);
    if (DIM == 3) q.z = add_mul(q.z, dt, v.z);
    D.posh[s] = q;
    if (!MIXED && P.has_solid && mt.x > 100) {
        const Material* m = find_mat(P, mt.x);
        if (m) {
            double* S = D.shear + (int64_t)s * DIM * DIM;
            const double* Sp = D.shear_prev + (int64_t)s * DIM * DIM;
            const double* dS = D.dsdt + (int64_t)s * DIM * DIM;
            for (int k = 0; k < DIM * DIM; ++k) S[k] = add_mul(first ? S[k] : Sp[k], fac, dS[k]);
            if (m->sigma_y > 0) {
                double J2 = 0.0;
                for (int k = 0; k < DIM * DIM; ++k) J2 = add_mul(J2, S[k], S[k]);
                double sc = fmin(1.0, sqrt(((m->sigma_y * m->sigma_y) / 3) / J2));
                for (int k = 0; k < DIM * DIM; ++k) S[k] = S[k] * sc;
            }
        }
    }
    if (!isfinite(q.x) || !isfinite(q.y) || !isfinite(q.z) || !isfinite(v.x) || !finite_rho)
        atomicOr(P.err, kErrNonFinite);
}

template <int DIM, bool MIXED, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_correct(Params P0, Dev D, int first_arg, RunMap M, int kstep) {
    int lb;
    const Params& P = block_run<BATCH>(P0, M, blockIdx.x, lb);
    const bool first = BATCH ? (P.step0 + kstep == 1) : first_arg != 0;
    const int l = lb * kBlock + threadIdx.x;
    const int s = P.s0 + l;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (l < P.n) {
        int4 mt = D.meta[s];
        double4 q = D.posh[s];
        if (mt.w < P.n_real) correct_one<DIM, MIXED>(P, D, s, mt, first, q);
        mn[0] = mx[0] = q.x; mn[1] = mx[1] = q.y; mn[2] = mx[2] = q.z;
    }
    bbox_block_reduce(mn, mx, P.partial, lb);
}

// Neighbour statistics of the last search (print_statistics, nnps.f90:377-416):
// out[0] total pairs, out[1] max, out[2] min, out[3] targets with none. A
// fixed grid of kStatsBlocks blocks strides over the particles and each block
// adds its results with one atomic per word: same-address atomics serialise
// in L2, so there are few of them.
constexpr int kStatsBlocks = 256;

__global__ __launch_bounds__(kBlock) void k_stats(Params P, Dev D, const int* counts,
                                                  unsigned long long* out) {
    unsigned long long tot = 0, mx = 0, mn = ~0ull, zero = 0;
    for (int s = blockIdx.x * kBlock + threadIdx.x; s < P.n; s += gridDim.x * kBlock) {
        const int4 mt = D.meta[s];
        if (mt.w < P.n_real) {
            const unsigned long long c = (mt.y == -1) ? 0 : (unsigned long long)counts[s];
            tot += c;
            mx = c > mx ? c : mx;
            mn = c < mn ? c : mn;
            zero += c == 0;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        tot += __shfl_xor(tot, o);
        zero += __shfl_xor(zero, o);
        unsigned long long a = __shfl_xor(mx, o), b = __shfl_xor(mn, o);
        mx = a > mx ? a : mx;
        mn = b < mn ? b : mn;
    }
    constexpr int kWaves = kBlock / 64;
    __shared__ unsigned long long red[4][kWaves];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = tot;
        red[1][w] = mx;
        red[2][w] = mn;
        red[3][w] = zero;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kWaves; ++k) {
            tot += red[0][k];
            mx = red[1][k] > mx ? red[1][k] : mx;
            mn = red[2][k] < mn ? red[2][k] : mn;
            zero += red[3][k];
        }
        atomicAdd(&out[0], tot);
        atomicMax(&out[1], mx);
        atomicMin(&out[2], mn);
        atomicAdd(&out[3], zero);
    }
}

// --------------------------------------------------------------------- host

namespace {

const char* kFieldNames[SPHGPU_F_NFIELDS] = {"x", "v", "rho", "p", "e", "mass", "h", "c",
                                             "disp", "stress", "shear", "dvdt", "drhodt",
                                             "dedt", "dsdt", "divx", "divv", "eta"};

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

// One run's particles in slot order with its parameters and grid plan: what
// an upload copies to the device. Slots are ordered by our linear cell key (x
// fastest): neighbouring lanes then hold x-neighbouring particles, so the k-th
// neighbours of a wavefront's lanes sit in consecutive slots and the gathers
// coalesce. A context keeps its HostRun between uploads, so a host that
// uploads every step (the drop-in single_step) reuses these buffers.
// A host array in slot order for an upload. A context keeps its arrays
// pinned (hipHostMalloc, grown on demand and reused by later uploads), so
// their copies to the device are DMA from pinned memory; a HostRun that does
// not ask for that, or that cannot get pinned memory, keeps them pageable.
template <typename T>
struct HostArr {
    bool want_pinned = false;
    T* pin = nullptr;
    size_t pin_n = 0, n_ = 0;
    std::vector<T> v;
    HostArr() = default;
    HostArr(const HostArr&) = delete;
    HostArr& operator=(const HostArr&) = delete;
    ~HostArr() { if (pin) hipHostFree(pin); }
    void resize(size_t n) {
        n_ = n;
        if (want_pinned) {
            if (n > pin_n) {
                if (pin) hipHostFree(pin);
                pin = nullptr;
                pin_n = 0;
                void* q = nullptr;
                if (hipHostMalloc(&q, n * sizeof(T), hipHostMallocDefault) == hipSuccess) {
                    pin = static_cast<T*>(q);
                    pin_n = n;
                } else {
                    (void)hipGetLastError();   // pageable from here on
                }
            }
            if (pin) return;
        }
        v.resize(n);
    }
    T* data() { return pin ? pin : v.data(); }
    const T* data() const { return pin ? pin : v.data(); }
    T& operator[](size_t i) { return data()[i]; }
    const T& operator[](size_t i) const { return data()[i]; }
    size_t size() const { return n_; }
    bool empty() const { return n_ == 0; }
    void clear() { n_ = 0; v.clear(); }
    const T* begin() const { return data(); }
    const T* end() const { return data() + n_; }
};

struct HostRun {
    Params P{};
    int64_t n = 0, n_real = 0, n_dummy = 0;
    bool fast = false;
    HostArr<double4> posh, vm;
    HostArr<double> rho, e, c, p, shear;   // shear empty when the host passes none
    HostArr<int4> meta;
    HostArr<float> drho;   // mixed precision: rho - rho0
    void pin_arrays() {
        posh.want_pinned = vm.want_pinned = rho.want_pinned = e.want_pinned = true;
        c.want_pinned = p.want_pinned = shear.want_pinned = meta.want_pinned = drho.want_pinned = true;
    }
    std::vector<uint32_t> sort_key, sort_start;   // counting-sort work space
    std::vector<int64_t> order;                   // slot -> particle index
    std::vector<int32_t> type_idx;                // type by particle index
    std::vector<int> sidx;                        // slots of the real solid targets
};

// Host staging for downloads: pinned (hipHostMalloc, grown on demand, kept
// by the context) or pageable.
struct HostStage {
    bool pinned = false;
    void* p = nullptr;
    size_t bytes = 0;
    std::vector<char> v;
    void* get(size_t need) {
        if (!pinned) {
            if (v.size() < need) v.resize(need);
            return v.data();
        }
        if (need > bytes) {
            release();
            if (hipHostMalloc(&p, need, hipHostMallocDefault) != hipSuccess) {
                // no pinned memory to be had: stage in pageable memory from now on
                p = nullptr;
                pinned = false;
                (void)hipGetLastError();
                if (v.size() < need) v.resize(need);
                return v.data();
            }
            bytes = need;
        }
        return p;
    }
    void release() {
        if (p) hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct sphgpu_ctx {
    sphgpu_config cfg{};
    Params P{};
    int64_t capacity = 0;
    int64_t n = 0, n_real = 0, n_dummy = 0;
    int dim = 2;
    hipStream_t stream = nullptr;
    // the general force pass of solid targets runs beside that of fluid
    // targets: a second stream, forked from and joined to `stream` by events
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    Dev D{};
    std::vector<void*> allocs;
    std::vector<size_t> alloc_bytes;   // size of each allocation (checkpoints copy them whole)
    int64_t ncell_alloc = 0;
    int64_t nbr_alloc = 0;
    int64_t vnbr_alloc = 0;
    int nblk = 0;
    int64_t step = 1;
    bool has_state = false;
    hipEvent_t ev[4]{};
    double last_advance_ms = 0, last_force_ms = 0;
    int64_t last_force_launches = 0;
    bool time_force = false;
    bool force_general = false;   // testing switch: use the general force kernel
    bool use_verlet = true;       // candidate lists (SPHGPU_NO_VERLET=1: exact lists by a cell sweep every step)
    bool fused_done = false;      // a step has run since the upload: the candidate list is valid
    bool use_graph = true;        // replay captured step sequences (SPHGPU_NO_GRAPH=1: eager)
    bool separate_filter = false; // testing switch: keep the candidate filter a pass of its own
    hipGraphExec_t graph = nullptr;   // kGraphSteps fast-path steps, valid until the next upload
    std::string err;
    int* h_err = nullptr;   // pinned
    // device buffers of the previous upload, reused by the next one when a
    // buffer of the same size is asked for (freed at the end of that upload)
    std::vector<void*> spare;
    std::vector<size_t> spare_bytes;
    HostRun up;              // the last upload in slot order
    bool up_meta = false;    // up.type_idx holds the uploaded particles' types
    HostStage dstage{true};  // pinned staging of downloads
    void* dorder = nullptr;  // device scratch: a downloaded field in particle order
};

static int fail(sphgpu_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                     \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(ctx, SPHGPU_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
static int dalloc(sphgpu_ctx* c, T** p, size_t count) {
    void* q = nullptr;
    size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    for (size_t k = 0; k < c->spare.size(); ++k)
        if (c->spare_bytes[k] == bytes) {
            q = c->spare[k];
            c->spare.erase(c->spare.begin() + k);
            c->spare_bytes.erase(c->spare_bytes.begin() + k);
            break;
        }
    if (!q) HIPCHK(c, hipMalloc(&q, bytes));
    HIPCHK(c, hipMemsetAsync(q, 0, bytes, c->stream));
    c->allocs.push_back(q);
    c->alloc_bytes.push_back(bytes);
    *p = static_cast<T*>(q);
    return 0;
}

static void drop_spare(sphgpu_ctx* c) {
    for (void* q : c->spare) hipFree(q);
    c->spare.clear();
    c->spare_bytes.clear();
}

// Every allocation becomes a spare for the allocations that follow (which
// zero what they take, as fresh ones are zeroed).
static void recycle_all(sphgpu_ctx* c) {
    c->spare.insert(c->spare.end(), c->allocs.begin(), c->allocs.end());
    c->spare_bytes.insert(c->spare_bytes.end(), c->alloc_bytes.begin(), c->alloc_bytes.end());
    c->allocs.clear();
    c->alloc_bytes.clear();
    c->D = Dev{};
    c->ncell_alloc = 0;
    c->nbr_alloc = 0;
    c->vnbr_alloc = 0;
}

static void free_all(sphgpu_ctx* c) {
    recycle_all(c);
    drop_spare(c);
}

static int validate_config(const sphgpu_config* cfg, std::string* why) {
    if (cfg->dim != 2 && cfg->dim != 3) { *why = "dim must be 2 or 3"; return SPHGPU_E_BAD_CONFIG; }
    if (cfg->skf < 1 || cfg->skf > 3) { *why = "skf must be 1, 2 or 3"; return SPHGPU_E_BAD_CONFIG; }
    if (cfg->pa_sph == 4) {
        // sum_density_dsph (density.f90:147-221) carries `ratio` and `s` from
        // one particle to the next and reads densities it has already
        // overwritten: its result depends on the particle order by design
        *why = "SPH = 4 (DSPH) is an order-dependent serial scheme; SPH = 1, 2, 3 are supported";
        return SPHGPU_E_BAD_CONFIG;
    }
    if (cfg->pa_sph < 1 || cfg->pa_sph > 3) { *why = "SPH must be 1, 2 or 3"; return SPHGPU_E_BAD_CONFIG; }
    if (cfg->pa_sph == 3 && cfg->summation_density) {
        *why = "SPH = 3 with summation density: the reference computes the Riemann rate and never applies it";
        return SPHGPU_E_BAD_CONFIG;
    }
    if (cfg->artificial_stress && cfg->dim != 2) {
        *why = "artificial stress is 2D only (artificialTerm.f90:94-95)";
        return SPHGPU_E_BAD_CONFIG;
    }
    if (cfg->n_materials < 0 || cfg->n_materials > SPHGPU_MAX_MATERIALS) {
        *why = "n_materials out of range";
        return SPHGPU_E_BAD_CONFIG;
    }
    if (cfg->sle < 0 || cfg->sle > 2) { *why = "sle must be 0, 1 or 2"; return SPHGPU_E_BAD_CONFIG; }
    if (cfg->precision != SPHGPU_FP64 && cfg->precision != SPHGPU_MIXED) {
        *why = "unknown precision";
        return SPHGPU_E_BAD_CONFIG;
    }
    if (!(cfg->dt > 0)) { *why = "dt must be positive"; return SPHGPU_E_BAD_CONFIG; }
    return 0;
}

extern "C" int sphgpu_abi_version(void) { return SPHGPU_ABI_VERSION; }

// Message of the last failed sphgpu_create on this thread (there is no
// context to hold it); sphgpu_last_error(NULL) returns it.
static thread_local std::string g_create_error;

static int create_fail(sphgpu_ctx* c, int rc, const std::string& why) {
    g_create_error = "sphgpu_create: " + why;
    if (c) {
        for (auto& e : c->ev)
            if (e) hipEventDestroy(e);
        if (c->stream) hipStreamDestroy(c->stream);
        delete c;
    }
    return rc;
}

static std::string hip_why(const char* call, hipError_t e) {
    return std::string(call) + " failed: " + hipGetErrorString(e);
}

extern "C" int sphgpu_create(const sphgpu_config* cfg, int64_t capacity, sphgpu_ctx** out) {
    if (!out) return create_fail(nullptr, SPHGPU_E_ARG, "out is NULL");
    *out = nullptr;
    if (!cfg || capacity <= 0) return create_fail(nullptr, SPHGPU_E_ARG, "NULL config or capacity <= 0");
    std::string why;
    int rc = validate_config(cfg, &why);
    if (rc) return create_fail(nullptr, rc, why);
    sphgpu_ctx* c = new (std::nothrow) sphgpu_ctx();
    if (!c) return create_fail(nullptr, SPHGPU_E_ARG, "out of host memory");
    c->up.pin_arrays();   // uploads stage in pinned memory kept by the context
    c->cfg = *cfg;
    c->dim = cfg->dim;
    c->capacity = capacity;
    const char* fg = getenv("SPHGPU_FORCE_GENERAL");
    c->force_general = fg && fg[0] == '1';
    const char* nv = getenv("SPHGPU_NO_VERLET");
    c->use_verlet = !(nv && nv[0] == '1');
    const char* ng = getenv("SPHGPU_NO_GRAPH");
    c->use_graph = !(ng && ng[0] == '1');
    const char* sf = getenv("SPHGPU_SEPARATE_FILTER");
    c->separate_filter = sf && sf[0] == '1';
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        c->stream = nullptr;
        return create_fail(c, SPHGPU_E_HIP, hip_why("hipStreamCreateWithFlags", e));
    }
    for (auto& ev : c->ev) {
        e = hipEventCreate(&ev);
        if (e != hipSuccess) {
            ev = nullptr;
            return create_fail(c, SPHGPU_E_HIP, hip_why("hipEventCreate", e));
        }
    }
    e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e != hipSuccess) {
        c->side = nullptr;
        return create_fail(c, SPHGPU_E_HIP, hip_why("hipStreamCreateWithFlags", e));
    }
    for (hipEvent_t* ev : {&c->fork, &c->join}) {
        e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            *ev = nullptr;
            return create_fail(c, SPHGPU_E_HIP, hip_why("hipEventCreateWithFlags", e));
        }
    }
    e = hipHostMalloc((void**)&c->h_err, sizeof(int) * 4, hipHostMallocDefault);
    if (e != hipSuccess) {
        c->h_err = nullptr;
        return create_fail(c, SPHGPU_E_HIP, hip_why("hipHostMalloc", e));
    }
    *out = c;
    return 0;
}

static void fill_params_cfg(Params& P, const sphgpu_config& g, int64_t n, int64_t n_real) {
    P.dim = g.dim;
    P.skf = g.skf;
    P.scale_k = (g.skf == 1) ? 2 : 3;   // asymmetricSearch.f90:178-185
    P.n = (int)n;
    P.n_real = (int)n_real;
    P.nmat = g.n_materials;
    P.summation_density = g.summation_density;
    P.normalize_density = g.normalize_density;
    P.viscosity = g.viscosity;
    P.gravity = g.gravity;
    P.gravity_mode = g.gravity_mode;
    P.external_force = g.external_force;
    P.arti_visc = g.artificial_viscosity;
    P.arti_stress = g.artificial_stress;
    P.kgc = g.kernel_correction;
    P.g = g.gravity_g;
    P.av_alpha = g.av_alpha; P.av_beta = g.av_beta; P.av_psi = g.av_psi;
    P.lj_factor = g.lj_factor; P.lj_r0 = g.lj_r0; P.lj_p1 = g.lj_p1; P.lj_p2 = g.lj_p2;
    P.pen_factor = g.pen_factor; P.pen_n1 = g.pen_n1; P.pen_n2 = g.pen_n2;
    P.couple_delta = g.couple_delta;
    P.as_lambda = g.as_lambda; P.as_delta = g.as_delta; P.as_exponent = g.as_exponent;
    P.dt = g.dt;
    P.mixed = g.precision == SPHGPU_MIXED;
    P.pair32 = 0;
    P.visc_skip_solid = g.fluid_visc_skip_solid != 0;
    P.sle = g.sle;
    P.pa_sph = g.pa_sph;
}

// The fused fluid force kernel covers the dam-break physics set; anything
// else (physical viscosity, solids, kernel correction, summation density,
// artificial stress, Gauss/quintic kernels) takes the general kernels.
static bool fast_path_ok(const sphgpu_config& g, const Params& P, bool force_general) {
    return g.skf == 1 && !g.viscosity && !P.has_solid && !g.kernel_correction &&
           !g.summation_density && !g.artificial_stress && g.sle == 0 && g.pa_sph == 2 &&
           !force_general;
}

static bool fluid_fast_path(const sphgpu_ctx* c) { return fast_path_ok(c->cfg, c->P, c->force_general); }

// Our search grid: cells of cutoff/win (slightly enlarged) over the upload
// bounding box plus a margin; keys outside are clamped to the border cells,
// which keeps every neighbour within the +-win sweep.
// Tuning of a run: cells per cutoff, the candidate skin (fraction of the
// cutoff) and the edges of the slot-order tiles in cells (y and z, and x);
// 0 takes the path's default. None changes which pairs interact; all change
// the speed.
struct Tuning {
    int win;
    double skin_frac;
    int tile;
    int tile_x;
};

// Default skins: the fluid fast path's fused pass tests every candidate with
// full pair arithmetic and is rebuilt rarely; the general path's filter is a
// pass of its own whose cost is one gather per candidate, so a thinner skin
// (more rebuilds, fewer candidates) is faster there.
constexpr double kSkinFast = 0.15, kSkinGeneral = 0.07;
// Default slot-order tiles in 3D (2D keeps whole rows and planes: its
// neighbours are a few rows apart). The fast path's fused pass gains most
// when the lanes of a wave are consecutive cells of one x-row, whose k-th
// candidates sit side by side: its tiles keep whole x-rows and bound y and z
// (8 cells). The general path's pair passes gather more bytes per pair and
// gain more from cubic tiles (64 cells), whose neighbour records stay in an
// XCD's L2. Both measured on the benchmark dam break and the elastic gate.
constexpr int kTileFast3 = 8, kTileGeneral3 = 64;
constexpr int kTileWhole = 1 << 30;

static Tuning env_tuning() {
    const char* wenv = getenv("SPHGPU_WIN");
    const char* senv = getenv("SPHGPU_SKIN");
    const char* tenv = getenv("SPHGPU_TILE");
    Tuning t;
    t.win = (wenv && wenv[0] >= '1' && wenv[0] <= '4') ? wenv[0] - '0' : kWinDefault;
    t.skin_frac = senv ? atof(senv) : 0.0;
    t.tile = tenv ? std::max(0, atoi(tenv)) : 0;
    const char* txenv = getenv("SPHGPU_TILE_X");
    t.tile_x = txenv ? std::max(0, atoi(txenv)) : 0;
    return t;
}

constexpr double kSleHeadroom = 1.5;

// f(lo, hi) over [0, n) split across up to 16 host threads (one thread for
// small n); each call owns its range, so the result does not depend on the
// split.
template <typename F>
static void host_parallel(int64_t n, F&& f) {
    const int64_t t = std::min<int64_t>(16, std::max(1u, std::thread::hardware_concurrency()));
    if (n < (1 << 16) || t == 1) {
        f((int64_t)0, n);
        return;
    }
    const int64_t chunk = (n + t - 1) / t;
    std::vector<std::thread> pool;
    for (int64_t k = 1; k < t; ++k) {
        const int64_t lo = k * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) continue;
        try {
            pool.emplace_back([&f, lo, hi] { f(lo, hi); });
        } catch (...) {   // no thread to be had: this range runs here
            f(lo, hi);
        }
    }
    f((int64_t)0, std::min(n, chunk));
    for (auto& th : pool) th.join();
}

// host_parallel with each range's number: f(r, lo, hi) for the ranges r =
// 0..R-1 of its split; returns R.
template <typename F>
static int64_t host_ranges(int64_t n, F&& f) {
    if (n <= 0) return 0;
    const int64_t t = n < (1 << 16) ? 1 : std::min<int64_t>(16, std::max(1u, std::thread::hardware_concurrency()));
    const int64_t chunk = (n + t - 1) / t;
    host_parallel(n, [&](int64_t lo, int64_t hi) { f(lo / chunk, lo, hi); });
    return (n + chunk - 1) / chunk;
}

// f(k) for k = 0..m-1, each on a host thread of its own (the calling thread
// takes k = 0, and any k whose thread cannot be created).
template <typename F>
static void host_tasks(int64_t m, F&& f) {
    std::vector<std::thread> pool;
    for (int64_t k = 1; k < m; ++k) {
        try {
            pool.emplace_back([&f, k] { f(k); });
        } catch (...) {
            f(k);
        }
    }
    if (m > 0) f(0);
    for (auto& th : pool) th.join();
}

static void plan_grid_params(Params& P, int dim, bool verlet, int64_t n, const double* x,
                             const double* h, Tuning tune, double h_next_max = 0.0) {
    double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0}, hmax = 0;
    for (int d = 0; d < dim; ++d) { mn[d] = INFINITY; mx[d] = -INFINITY; }
    {
        // per-range extremes, merged (min and max do not depend on the order)
        std::vector<double> ext(16 * 7, 0.0);
        const int64_t nr = host_ranges(n, [&](int64_t r, int64_t lo, int64_t hi) {
            double a[3] = {INFINITY, INFINITY, INFINITY}, b[3] = {-INFINITY, -INFINITY, -INFINITY}, hm = 0;
            for (int64_t i = lo; i < hi; ++i) {
                for (int d = 0; d < dim; ++d) {
                    double v = x[i * dim + d];
                    if (v < a[d]) a[d] = v;
                    if (v > b[d]) b[d] = v;
                }
                hm = std::max(hm, h[i]);
            }
            for (int d = 0; d < 3; ++d) { ext[r * 7 + d] = a[d]; ext[r * 7 + 3 + d] = b[d]; }
            ext[r * 7 + 6] = hm;
        });
        for (int64_t r = 0; r < nr; ++r) {
            for (int d = 0; d < dim; ++d) {
                mn[d] = std::min(mn[d], ext[r * 7 + d]);
                mx[d] = std::max(mx[d], ext[r * 7 + 3 + d]);
            }
            hmax = std::max(hmax, ext[r * 7 + 6]);
        }
    }
    // an evolving h gets headroom: the grid, the window test and the
    // neighbour capacity hold for h up to kSleHeadroom times the larger of
    // its upload value and its first update (sle 1 resets h from density)
    P.hmax_upload = hmax;
    if (P.sle) hmax = std::max(hmax, h_next_max) * kSleHeadroom;
    P.win = tune.win;
    const double cut_max = hmax * P.scale_k;
    P.skin = verlet ? tune.skin_frac * cut_max : 0.0;
    P.hmax = hmax;
    const double half = 0.5 * P.skin * (1.0 - 1e-9);
    P.disp_thr2 = half * half;
    P.disp_half = half;
    double cs = (cut_max + P.skin) / P.win * (1.0 + 1e-6);
    if (!(cs > 0)) cs = 1.0;
    double lo[3], ext[3];
    for (int d = 0; d < dim; ++d) {
        double span = mx[d] - mn[d];
        double margin = 0.05 * span + 2 * cs;
        lo[d] = mn[d] - margin;
        ext[d] = span + 2 * margin;
    }
    const int64_t budget = std::max<int64_t>(8 * n, 4096);
    // per-axis limits of the packed cell coordinates (11/11/9 bits in 3D, 16/15
    // in 2D; bit 31 marks a particle outside the grid)
    const int64_t lim[3] = {dim == 3 ? 2047 : 65535, dim == 3 ? 2047 : 32767, 511};
    for (;;) {
        // the cell size is a float value, so FP32 cell offsets reconstruct pair
        // separations exactly in the cell part
        cs = (double)(float)cs;
        int64_t cells = 1;
        bool ok = true;
        for (int d = 0; d < dim; ++d) {
            int64_t nd = (int64_t)std::ceil(ext[d] / cs) + 1;
            cells *= nd;
            ok &= nd <= lim[d];
        }
        if (ok && cells <= budget && cells < (1LL << 31) - 2) break;
        cs *= 1.25;
    }
    for (int d = 0; d < 3; ++d) { P.org[d] = 0; P.nc[d] = 1; }
    for (int d = 0; d < dim; ++d) {
        P.org[d] = lo[d];
        P.nc[d] = (int)std::ceil(ext[d] / cs) + 1;
    }
    P.inv_cs = 1.0 / cs;
    P.cs = cs;
    P.ncell = (int64_t)P.nc[0] * P.nc[1] * P.nc[2];
}

static void plan_grid(sphgpu_ctx* c, const double* x, const double* h) {
    plan_grid_params(c->P, c->dim, c->use_verlet, c->n, x, h, env_tuning());
}

template <int DIM>
static void launch_sort(sphgpu_ctx* c, Gate G = Gate{nullptr}, int with_ref = 1) {
    Params& P = c->P;
    const int nb = (int)((P.n + kBlock - 1) / kBlock);
    const int nbs = std::min(nb, kStrideBlocks);
    if (P.n <= kSmallSortSlots && P.ncell + 1 <= kSmallSortCells) {
        hipLaunchKernelGGL(k_sort_small<DIM>, dim3(nbs), dim3(kBlock), 0, c->stream, P, c->D, G, with_ref,
                           (int)P.ncell + 1);
        return;
    }
    hipLaunchKernelGGL((k_keys<DIM, false>), dim3(nbs), dim3(kBlock), 0, c->stream, P, c->D, G, with_ref, RunMap{});
    const int m = (int)P.ncell + 1;
    const int sb = (m + kScanTile - 1) / kScanTile;
    const int own = sb <= kOwnPrefixTiles;
    hipLaunchKernelGGL(k_scan1, dim3(sb), dim3(kBlock), 0, c->stream, c->D.count, m, c->D.bsum, G);
    if (!own) hipLaunchKernelGGL(k_scan2, dim3(1), dim3(kBlock), 0, c->stream, c->D.bsum, sb, G);
    hipLaunchKernelGGL(k_scan3, dim3(sb), dim3(kBlock), 0, c->stream, c->D.count, m, c->D.bsum,
                       c->D.start, own, G);
    hipLaunchKernelGGL(k_scatter<false>, dim3(nbs), dim3(kBlock), 0, c->stream, P, c->D, G, RunMap{});
    hipLaunchKernelGGL(k_cellsort<false>, dim3(std::min((m + kBlock - 1) / kBlock, kStrideBlocks)), dim3(kBlock),
                       0, c->stream, P, c->D, m, G, RunMap{});
}

static Gate fast_gate(const sphgpu_ctx* c) { return Gate{c->D.vflag}; }

// The fluid fast path. Head: reference grid, predict + EOS + displacement
// check. Tail: candidate-list rebuild when the gate is open, fused force pass.
template <int DIM>
static void launch_fast_head(sphgpu_ctx* c, int mode) {
    Params& P = c->P;
    const int nb = (int)((P.n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL((k_grid<DIM, false>), dim3(1), dim3(kGridThreads), 0, c->stream, P, c->D, RunMap{});
    if (P.mixed)
        hipLaunchKernelGGL((k_point_fast<DIM, true, false>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D, mode, RunMap{});
    else
        hipLaunchKernelGGL((k_point_fast<DIM, false, false>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D, mode, RunMap{});
}

template <int DIM>
static void launch_fast_tail(sphgpu_ctx* c, int count_only) {
    Params& P = c->P;
    const int nb = (int)((P.n + kBlock - 1) / kBlock);
    const Gate G = fast_gate(c);
    launch_sort<DIM>(c, G, 2);
    hipLaunchKernelGGL((k_verlet_build<DIM, false>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D, G, count_only, RunMap{});
    if (count_only) return;
    if (c->time_force) hipEventRecord(c->ev[2], c->stream);
    if (P.mixed && P.pair_ext)
        hipLaunchKernelGGL((k_force_fused<DIM, true, true, false>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D, RunMap{});
    else if (P.mixed)
        hipLaunchKernelGGL((k_force_fused<DIM, true, false, false>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D, RunMap{});
    else if (P.pair_ext)
        hipLaunchKernelGGL((k_force_fused<DIM, false, true, false>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D, RunMap{});
    else
        hipLaunchKernelGGL((k_force_fused<DIM, false, false, false>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D, RunMap{});
    if (c->time_force) hipEventRecord(c->ev[3], c->stream);
}

template <int DIM>
static void launch_fast(sphgpu_ctx* c, int mode, int count_only = 0) {
    launch_fast_head<DIM>(c, mode);
    launch_fast_tail<DIM>(c, count_only);
}

// One step of the fast path that is followed by another: list rebuild if due,
// fused force pass, correct of this step fused with the predict + EOS of the
// next, and the next step's reference grid and cells.
template <int DIM>
static void fast_body_step(sphgpu_ctx* c, bool timed) {
    Params& P = c->P;
    const int nb = (int)((P.n + kBlock - 1) / kBlock);
    c->time_force = timed;
    launch_fast_tail<DIM>(c, 0);
    const int first = c->step == 1;
    if (P.mixed)
        hipLaunchKernelGGL((k_correct_predict<DIM, true, false>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D, first, RunMap{}, 0);
    else
        hipLaunchKernelGGL((k_correct_predict<DIM, false, false>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D, first, RunMap{}, 0);
    hipLaunchKernelGGL((k_grid<DIM, false>), dim3(1), dim3(kGridThreads), 0, c->stream, P, c->D, RunMap{});
    c->time_force = false;
}

// The last step of a call: force pass and a plain correct.
template <int DIM>
static void fast_last_step(sphgpu_ctx* c, bool timed) {
    Params& P = c->P;
    const int nb = (int)((P.n + kBlock - 1) / kBlock);
    c->time_force = timed;
    launch_fast_tail<DIM>(c, 0);
    const int first = c->step == 1;
    if (P.mixed)
        hipLaunchKernelGGL((k_correct<DIM, true, false>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D, first, RunMap{}, 0);
    else
        hipLaunchKernelGGL((k_correct<DIM, false, false>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D, first, RunMap{}, 0);
    c->time_force = false;
}

// Steps per captured graph of the fast path.
constexpr int kGraphSteps = 10;

static void drop_graph(sphgpu_ctx* c) {
    if (c->graph) hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
}

// Capture kGraphSteps steps of `body` (kernel nodes only: no memsets, no
// events). The kernels take the uploaded Params and device arrays by value,
// and a step after the first has the same arguments every time, so the graph
// holds until the next upload. Capturing launches nothing, so it is done on
// the first advance after an upload whatever its length (the steps are
// captured as steps after the first), and no later call pays for it.
template <typename F>
static int capture_steps(sphgpu_ctx* c, F&& body) {
    const int64_t step = c->step;
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->step = std::max<int64_t>(step, 2);
    c->time_force = false;
    for (int k = 0; k < kGraphSteps; ++k) {
        body();
        c->step++;
    }
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->step = step;
    if (e != hipSuccess) return fail(c, SPHGPU_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        c->graph = nullptr;
        return fail(c, SPHGPU_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    return 0;
}

// nsteps leap-frog steps on the fast path: predict + EOS of the first step,
// then nsteps - 1 body steps (whole groups of kGraphSteps replay a captured
// graph once the first step is past) and a closing step whose force kernel
// is timed.
template <int DIM>
static int advance_fast(sphgpu_ctx* c, int nsteps) {
    if (c->use_graph && !c->graph) {
        int rc = capture_steps(c, [c] { fast_body_step<DIM>(c, false); });
        if (rc) return rc;
    }
    launch_fast_head<DIM>(c, c->step == 1 ? 1 : 0);
    int body = nsteps - 1;
    if (c->use_graph && c->step > 1 && body >= kGraphSteps) {
        const int reps = body / kGraphSteps;
        for (int r = 0; r < reps; ++r) {
            HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
            c->step += kGraphSteps;
        }
        body -= reps * kGraphSteps;
    }
    for (int k = 0; k < body; ++k) {
        fast_body_step<DIM>(c, false);
        c->step++;
    }
    fast_last_step<DIM>(c, true);
    c->step++;
    return 0;
}

static void force_rebuild(sphgpu_ctx* c) {
    hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(64), 0, c->stream, c->D.vflag);
}


// One evaluation of the derivatives (mode bits as k_point_build).
template <int DIM>
static void launch_point_build(sphgpu_ctx* c, int mode) {
    const int nb = (int)((c->P.n + kBlock - 1) / kBlock);
    if (c->P.mixed)
        hipLaunchKernelGGL((k_point_build<DIM, true>), dim3(nb), dim3(kBlock), 0, c->stream, c->P, c->D, mode);
    else
        hipLaunchKernelGGL((k_point_build<DIM, false>), dim3(nb), dim3(kBlock), 0, c->stream, c->P, c->D, mode);
}

// In mixed precision the candidate filter runs inside the strain pass when
// that is the step's first pair pass (no kernel gradient correction or
// summation density before it) and takes one lane per target; otherwise it is
// a pass of its own. (FP64 runs it alone: the fused FP64 pass holds too many
// registers to gain.) SPHGPU_SEPARATE_FILTER=1 keeps it alone (testing).
static bool strain_filters(const sphgpu_ctx* c) {
    const Params& P = c->P;
    return c->use_verlet && !c->separate_filter && P.pair32 && !fluid_fast_path(c) && !P.kgc &&
           !P.summation_density && (P.viscosity || P.has_solid) && P.n_real > kSmallRun;
}

// The general path's pair passes before the force pass: kernel gradient
// correction, summation density, strain rates, with LPT lanes per target.
template <int DIM, typename R, int LPT>
static void launch_pair_passes(sphgpu_ctx* c) {
    const Params& P = c->P;
    const int nb = (int)((P.n + kBlock - 1) / kBlock);
    const int nbl = (int)(((int64_t)P.n * LPT + kBlock - 1) / kBlock);
    if (P.kgc) hipLaunchKernelGGL((k_kgc<DIM, R, LPT>), dim3(nbl), dim3(kBlock), 0, c->stream, P, c->D);
    if (P.summation_density) {
        hipLaunchKernelGGL((k_sumdens<DIM, R, LPT>), dim3(nbl), dim3(kBlock), 0, c->stream, P, c->D);
        hipLaunchKernelGGL(k_apply_sumdens, dim3(nb), dim3(kBlock), 0, c->stream, P, c->D);
    }
    if ((P.viscosity || P.has_solid) && P.kgc)
        hipLaunchKernelGGL((k_strain<DIM, R, LPT, true>), dim3(nbl), dim3(kBlock), 0, c->stream, P, c->D);
    else if (LPT == 1 && strain_filters(c))
        hipLaunchKernelGGL((k_strain_filter<DIM, R>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D);
    else if (P.viscosity || P.has_solid)
        hipLaunchKernelGGL((k_strain<DIM, R, LPT, false>), dim3(nbl), dim3(kBlock), 0, c->stream, P, c->D);
}

// The general force pass: one launch per target class present (fluid, solid),
// each with the SPH = 2 form compiled alone when that is the run's form.
template <int DIM, typename R, int SPH>
static void launch_force_classes(sphgpu_ctx* c) {
    const Params& P = c->P;
    const int nb = (int)((P.n + kBlock - 1) / kBlock);
    if (P.n_real <= kSmallRun) {
        const int nbl = (int)(((int64_t)P.n * kLanesPerTarget + kBlock - 1) / kBlock);
        hipLaunchKernelGGL((k_force<DIM, R, kTgtAll, SPH, kLanesPerTarget, true>), dim3(nbl), dim3(kBlock), 0,
                           c->stream, P, c->D);
        return;
    }
    if (P.has_t100) {
        hipLaunchKernelGGL((k_force<DIM, R, kTgtAll, SPH, 1, true>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D);
        return;
    }
    // the solid targets are few: they take several lanes each, on the side
    // stream while the fluid pass runs
    const bool solids = P.n_sidx > 0;
    if (solids) {
        const int nbs = (int)(((int64_t)P.n_sidx * kLanesPerTarget + kBlock - 1) / kBlock);
        hipEventRecord(c->fork, c->stream);
        hipStreamWaitEvent(c->side, c->fork, 0);
        hipLaunchKernelGGL((k_force<DIM, R, kTgtSolid, SPH, kLanesPerTarget, true>), dim3(nbs), dim3(kBlock), 0,
                           c->side, P, c->D);
        hipEventRecord(c->join, c->side);
    }
    if (P.ext_pow)
        hipLaunchKernelGGL((k_force<DIM, R, kTgtFluid, SPH, 1, true>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D);
    else
        hipLaunchKernelGGL((k_force<DIM, R, kTgtFluid, SPH, 1, false>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D);
    if (solids) hipStreamWaitEvent(c->stream, c->join, 0);
}

template <int DIM, typename R>
static void launch_force_general(sphgpu_ctx* c) {
    if (c->P.pa_sph == 2) launch_force_classes<DIM, R, 2>(c);
    else launch_force_classes<DIM, R, 0>(c);
}

// Pair passes and force pass of one evaluation of the derivatives, after the
// search of any path but the fluid fast path's.
template <int DIM>
static void launch_physics(sphgpu_ctx* c) {
    Params& P = c->P;
    const int nb = (int)((P.n + kBlock - 1) / kBlock);
    const bool fast = fluid_fast_path(c);
    if (!fast) {
        // pair32: the pair sums of the general path in FP32 (mixed precision)
        if (P.n_real <= kSmallRun) {
            if (P.pair32) launch_pair_passes<DIM, float, kLanesPerTarget>(c);
            else launch_pair_passes<DIM, double, kLanesPerTarget>(c);
        } else {
            if (P.pair32) launch_pair_passes<DIM, float, 1>(c);
            else launch_pair_passes<DIM, double, 1>(c);
        }
        if (P.has_solid)
            hipLaunchKernelGGL(k_solid_point<DIM>, dim3(nb), dim3(kBlock), 0, c->stream, P, c->D);
    }
    if (c->time_force) hipEventRecord(c->ev[2], c->stream);
    if (P.mixed && P.pair_ext)
        hipLaunchKernelGGL((k_force_fluid<DIM, true, true>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D);
    else if (P.mixed)
        hipLaunchKernelGGL((k_force_fluid<DIM, true, false>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D);
    else if (fast && P.pair_ext)
        hipLaunchKernelGGL((k_force_fluid<DIM, false, true>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D);
    else if (fast)
        hipLaunchKernelGGL((k_force_fluid<DIM, false, false>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D);
    else if (P.pair32)
        launch_force_general<DIM, float>(c);
    else
        launch_force_general<DIM, double>(c);
    if (c->time_force) hipEventRecord(c->ev[3], c->stream);
    if (P.sle) hipLaunchKernelGGL(k_hsml<DIM>, dim3(nb), dim3(kBlock), 0, c->stream, P, c->D);
}

// The general path with a candidate list. Head: reference grid, predict +
// EOS + records + reference cells + displacement vote (mode bits as
// k_point_gen). Tail: the rebuild (cell sort and candidate list) when the
// gate is open, this step's exact list from the candidates, the pair and
// force passes. count_only: size the candidate list and stop.
template <int DIM>
static void gen_head(sphgpu_ctx* c, int mode) {
    const int nb = (int)((c->P.n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL((k_grid<DIM, false>), dim3(1), dim3(kGridThreads), 0, c->stream, c->P, c->D, RunMap{});
    hipLaunchKernelGGL(k_point_gen<DIM>, dim3(nb), dim3(kBlock), 0, c->stream, c->P, c->D, mode);
}

template <int DIM>
static void gen_tail(sphgpu_ctx* c, int count_only) {
    Params& P = c->P;
    const int nb = (int)((P.n + kBlock - 1) / kBlock);
    const Gate G = fast_gate(c);
    // the reference cells of this step ride on the key launch (with_ref 2)
    launch_sort<DIM>(c, G, 2);
    hipLaunchKernelGGL((k_verlet_build<DIM, false>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D, G, count_only,
                       RunMap{});
    if (count_only) return;
    if (strain_filters(c)) {
    } else if (P.n_real <= kSmallRun) {
        const int nbl = (int)(((int64_t)P.n * kLanesPerTarget + kBlock - 1) / kBlock);
        hipLaunchKernelGGL((k_filter<DIM, kLanesPerTarget>), dim3(nbl), dim3(kBlock), 0, c->stream, P, c->D);
    } else {
        hipLaunchKernelGGL((k_filter<DIM, 1>), dim3(nb), dim3(kBlock), 0, c->stream, P, c->D);
    }
    launch_physics<DIM>(c);
}

template <int DIM>
static void launch_derivs(sphgpu_ctx* c, int mode) {
    if (fluid_fast_path(c) && c->use_verlet) {
        launch_fast<DIM>(c, mode);
        return;
    }
    if (c->use_verlet) {
        gen_head<DIM>(c, mode);
        gen_tail<DIM>(c, 0);
        return;
    }
    hipLaunchKernelGGL((k_grid<DIM, false>), dim3(1), dim3(kGridThreads), 0, c->stream, c->P, c->D, RunMap{});
    launch_sort<DIM>(c);
    launch_point_build<DIM>(c, mode);
    launch_physics<DIM>(c);
}

// One step of the exact-list paths: search, physics, correct.
template <int DIM>
static void launch_step(sphgpu_ctx* c) {
    const int nb = (int)((c->P.n + kBlock - 1) / kBlock);
    int first = c->step == 1;
    launch_derivs<DIM>(c, first ? 1 : 0);
    if (c->P.mixed)
        hipLaunchKernelGGL((k_correct<DIM, true, false>), dim3(nb), dim3(kBlock), 0, c->stream, c->P, c->D, first, RunMap{}, 0);
    else
        hipLaunchKernelGGL((k_correct<DIM, false, false>), dim3(nb), dim3(kBlock), 0, c->stream, c->P, c->D, first, RunMap{}, 0);
}

// A step of the general path with a candidate list that is followed by
// another: its tail, the correct fused with the next step's predict, and the
// next step's reference grid (its reference cells come with the next tail's
// keys).
template <int DIM>
static void gen_body_step(sphgpu_ctx* c) {
    const int nb = (int)((c->P.n + kBlock - 1) / kBlock);
    const int first = c->step == 1;
    gen_tail<DIM>(c, 0);
    hipLaunchKernelGGL(k_correct_point_gen<DIM>, dim3(nb), dim3(kBlock), 0, c->stream, c->P, c->D, first);
    hipLaunchKernelGGL((k_grid<DIM, false>), dim3(1), dim3(kGridThreads), 0, c->stream, c->P, c->D, RunMap{});
}

// nsteps steps of the general path. With a candidate list: the head of the
// first step, nsteps - 1 body steps (whole groups of kGraphSteps replay a
// captured graph once the first step is past) and a closing step whose force
// pass is timed, as on the fast path. Exact lists: whole steps, groups of
// kGraphSteps replayed while more than one step is left.
template <int DIM>
static int advance_general(sphgpu_ctx* c, int nsteps) {
    if (nsteps <= 0) return 0;
    if (c->use_verlet) {
        if (c->use_graph && !c->graph) {
            int rc = capture_steps(c, [c] { gen_body_step<DIM>(c); });
            if (rc) return rc;
        }
        gen_head<DIM>(c, c->step == 1 ? 1 : 0);
        int body = nsteps - 1;
        if (c->use_graph && c->step > 1 && body >= kGraphSteps) {
            const int reps = body / kGraphSteps;
            for (int r = 0; r < reps; ++r) {
                HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
                c->step += kGraphSteps;
            }
            body -= reps * kGraphSteps;
        }
        for (int k = 0; k < body; ++k) {
            gen_body_step<DIM>(c);
            c->step++;
        }
        const int nb = (int)((c->P.n + kBlock - 1) / kBlock);
        const int first = c->step == 1;
        c->time_force = true;
        gen_tail<DIM>(c, 0);
        c->time_force = false;
        hipLaunchKernelGGL((k_correct<DIM, false, false>), dim3(nb), dim3(kBlock), 0, c->stream, c->P, c->D, first,
                           RunMap{}, 0);
        c->step++;
        return 0;
    }
    if (c->use_graph && !c->graph) {
        int rc = capture_steps(c, [c] { launch_step<DIM>(c); });
        if (rc) return rc;
    }
    for (int k = 0; k < nsteps;) {
        if (c->use_graph && c->step > 1 && nsteps - k > kGraphSteps) {
            HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
            c->step += kGraphSteps;
            k += kGraphSteps;
            continue;
        }
        c->time_force = (k == nsteps - 1);
        launch_step<DIM>(c);
        c->step++;
        ++k;
    }
    c->time_force = false;
    return 0;
}

static int check_err(sphgpu_ctx* c) {
    HIPCHK(c, hipMemcpyAsync(c->h_err, c->D.err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int e = c->h_err[0];
    if (e) {
        HIPCHK(c, hipMemsetAsync(c->D.err, 0, sizeof(int), c->stream));
        if (e & kErrNonFinite) return fail(c, SPHGPU_E_NONFINITE, "non-finite position, velocity or density");
        if (e & kErrOutOfRange) return fail(c, SPHGPU_E_OUT_OF_RANGE, "particle out of range");
        if (e & kErrPairCap)
            return fail(c, SPHGPU_E_PAIR_CAPACITY,
                        "neighbour list capacity exceeded: the lists are planned at upload for 1.5x its "
                        "largest neighbour count + 8, and this step's pair sums are incomplete "
                        "(upload the last good state again)");
        if (e & kErrHRange)
            return fail(c, SPHGPU_E_OUT_OF_RANGE,
                        "smoothing length grew past the planned search range (upload the state again)");
    }
    return 0;
}

static int alloc_state(sphgpu_ctx* c) {
    recycle_all(c);
    const int64_t N = c->capacity;
    const int dim = c->dim;
    const int64_t T = N * dim * dim;
    Dev& D = c->D;
    int rc = 0;
    rc |= dalloc(c, &D.posh, N); rc |= dalloc(c, &D.vm, N); rc |= dalloc(c, &D.rho, N);
    rc |= dalloc(c, &D.e, N); rc |= dalloc(c, &D.c, N); rc |= dalloc(c, &D.p, N);
    rc |= dalloc(c, &D.meta, N); rc |= dalloc(c, &D.dv, N);
    if (c->P.mixed) {
        rc |= dalloc(c, &D.dispf, N); rc |= dalloc(c, &D.xbf, N);
        rc |= dalloc(c, &D.ef, N); rc |= dalloc(c, &D.eprevf, N); rc |= dalloc(c, &D.divf, N);
    } else {
        rc |= dalloc(c, &D.disp, N); rc |= dalloc(c, &D.xb, N);
    }
    rc |= dalloc(c, &D.de, N); rc |= dalloc(c, &D.vprev, N); rc |= dalloc(c, &D.eprev, N);
    rc |= dalloc(c, &D.th, N); rc |= dalloc(c, &D.div, N); rc |= dalloc(c, &D.eta, N);
    rc |= dalloc(c, &D.tds, N); rc |= dalloc(c, &D.key, N); rc |= dalloc(c, &D.refcell, N);
    rc |= dalloc(c, &D.perm, N); rc |= dalloc(c, &D.ncnt, N);
    rc |= dalloc(c, &D.cpos, N);
    rc |= dalloc(c, &D.vcnt, N); rc |= dalloc(c, &D.ecnt, N);
    // FP32 position records: the mixed fast path's pair inputs, the pre-test
    // of every candidate path and the mixed general path's pair geometry
    if (c->P.mixed || c->use_verlet || c->P.pair32) rc |= dalloc(c, &D.posf, N);
    // the mixed general path gathers FP32 pair records and tensors
    if (c->P.pair32) {
        rc |= dalloc(c, &D.prf, N);
        if (c->cfg.viscosity || c->P.has_solid) rc |= dalloc(c, &D.Epf, T);
        if (c->P.has_solid) rc |= dalloc(c, &D.Spf, T);
    }
    rc |= dalloc(c, &D.vflag, 2);
    bool need_tensor = c->cfg.viscosity || c->P.has_solid;
    if (need_tensor) rc |= dalloc(c, &D.Ep, T);
    // S, stress and dS/dt always exist (zero for fluids) so they can be read back
    rc |= dalloc(c, &D.shear, T); rc |= dalloc(c, &D.stress, T);
    rc |= dalloc(c, &D.dsdt, T);
    if (c->P.has_solid) {
        rc |= dalloc(c, &D.Ed, T); rc |= dalloc(c, &D.rot, T);
        rc |= dalloc(c, &D.shear_prev, T); rc |= dalloc(c, &D.Sp, T);
        rc |= dalloc(c, &D.asR, N * 4);
        rc |= dalloc(c, &D.sidx, N);
    }
    if (c->cfg.kernel_correction) rc |= dalloc(c, &D.kgcA, T);
    if (c->P.mixed) {
        rc |= dalloc(c, &D.prf, N); rc |= dalloc(c, &D.drho, N);
        rc |= dalloc(c, &D.drho_prev, N); rc |= dalloc(c, &D.dvf, N); rc |= dalloc(c, &D.def, N);
    }
    if (c->cfg.summation_density) rc |= dalloc(c, &D.rho_new, N);
    if (c->cfg.pa_sph == 3) rc |= dalloc(c, &D.pold, N);
    int nb = (int)((N + kBlock - 1) / kBlock);
    rc |= dalloc(c, &D.partial, (size_t)nb * 6);
    rc |= dalloc(c, &D.grid, 1);
    rc |= dalloc(c, &D.err, 4);
    rc |= dalloc(c, &D.maxcnt, 4);
    rc |= dalloc(c, &D.mat, SPHGPU_MAX_MATERIALS);
    if (rc) return SPHGPU_E_HIP;
    Material tab[SPHGPU_MAX_MATERIALS];
    for (int k = 0; k < SPHGPU_MAX_MATERIALS; ++k) {
        const sphgpu_material& m = c->cfg.mat[k];
        tab[k] = Material{m.type, m.eos, m.rho0, m.K, m.gamma, m.eta, m.sigma_y};
    }
    HIPCHK(c, hipMemcpyAsync(D.mat, tab, sizeof(tab), hipMemcpyHostToDevice, c->stream));
    c->P.mat = D.mat;
    // a single run owns the whole arena
    c->P.s0 = 0;
    c->P.c0 = 0;
    c->P.nblk = (int)((c->n + kBlock - 1) / kBlock);
    c->P.step0 = 1;
    c->P.grid = D.grid;
    c->P.vflag = D.vflag;
    c->P.vany = D.vflag;
    c->P.err = D.err;
    c->P.maxcnt = D.maxcnt;
    c->P.partial = D.partial;
    return 0;
}

static int alloc_cells(sphgpu_ctx* c) {
    const int64_t m = c->P.ncell + 1;
    if (m <= c->ncell_alloc) return 0;
    Dev& D = c->D;
    int rc = 0;
    rc |= dalloc(c, &D.count, m + 1);
    rc |= dalloc(c, &D.start, m + 1);
    rc |= dalloc(c, &D.bsum, (m + kScanTile - 1) / kScanTile + 1);
    if (rc) return SPHGPU_E_HIP;
    c->ncell_alloc = m;
    return 0;
}

static int alloc_nbr(sphgpu_ctx* c, int cap) {
    int64_t need = vlist_entries(c->capacity, cap);
    if (need <= c->nbr_alloc) return 0;
    if (dalloc(c, &c->D.nbr, (size_t)need)) return SPHGPU_E_HIP;
    c->nbr_alloc = need;
    return 0;
}

// Testing switch SPHGPU_VERLET_CAP=k: at most k candidate rows per particle,
// so that particles past it take the overflow walk of their cell window.
static int test_verlet_cap(int cap) {
    const char* v = getenv("SPHGPU_VERLET_CAP");
    return v ? std::min(cap, std::max(1, atoi(v))) : cap;
}

template <int DIM>
static int size_neighbour_list(sphgpu_ctx* c) {
    // count-only passes on the uploaded state, then capacities with headroom
    const int nb = (int)((c->P.n + kBlock - 1) / kBlock);
    HIPCHK(c, hipMemsetAsync(c->D.maxcnt, 0, sizeof(int), c->stream));
    hipLaunchKernelGGL(k_bbox<false>, dim3(nb), dim3(kBlock), 0, c->stream, c->P, c->D, RunMap{});
    hipLaunchKernelGGL((k_grid<DIM, false>), dim3(1), dim3(kGridThreads), 0, c->stream, c->P, c->D, RunMap{});
    launch_sort<DIM>(c);
    launch_point_build<DIM>(c, 2 | 4);
    HIPCHK(c, hipMemcpyAsync(c->h_err + 1, c->D.maxcnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int e = check_err(c);
    if (e) return e;
    int mx = c->h_err[1];
    int cap = mx + std::max(8, mx / 2);
    // the count was taken with the uploaded h; an evolving h may grow to the
    // planned limit, and the neighbour count with its dim-th power
    if (c->P.sle && c->P.hmax_upload > 0)
        cap = (int)std::ceil(cap * std::pow(c->P.hmax / c->P.hmax_upload, c->dim));
    c->P.nbr_cap = cap;
    e = alloc_nbr(c, cap);
    if (e) return e;
    if (fluid_fast_path(c) && c->use_verlet) {
        HIPCHK(c, hipMemsetAsync(c->D.maxcnt, 0, sizeof(int), c->stream));
        force_rebuild(c);
        launch_fast<DIM>(c, 2, 1);
        HIPCHK(c, hipMemcpyAsync(c->h_err + 1, c->D.maxcnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        e = check_err(c);
        if (e) return e;
        int vx = c->h_err[1];
        int vcap = test_verlet_cap(vx + std::max(8, vx / 2));
        c->P.verlet_cap = vcap;
        const int64_t need = vlist_entries(c->capacity, vcap);
        if (need > c->vnbr_alloc) {
            if (dalloc(c, &c->D.vnbr, (size_t)need)) return SPHGPU_E_HIP;
            c->vnbr_alloc = need;
        }
    } else if (c->use_verlet) {
        HIPCHK(c, hipMemsetAsync(c->D.maxcnt, 0, sizeof(int), c->stream));
        force_rebuild(c);
        gen_head<DIM>(c, 2 | 4);
        gen_tail<DIM>(c, 1);
        HIPCHK(c, hipMemcpyAsync(c->h_err + 1, c->D.maxcnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        e = check_err(c);
        if (e) return e;
        int vx = c->h_err[1];
        int vcap = test_verlet_cap(vx + std::max(8, vx / 2));
        // an evolving h: the list is built for the planned largest h, whose
        // candidate count the sizing pass (at the uploaded h) has taken
        // already; the headroom covers the particles crowding in
        c->P.verlet_cap = vcap;
        const int64_t need = vlist_entries(c->capacity, vcap);
        if (need > c->vnbr_alloc) {
            if (dalloc(c, &c->D.vnbr, (size_t)need)) return SPHGPU_E_HIP;
            c->vnbr_alloc = need;
        }
    }
    return 0;
}


static int prepare_run(const sphgpu_config& cfg, int64_t n_real, int64_t n_dummy, const double* x,
                       const double* v, const double* mass, const double* rho, const double* p,
                       const double* e, const double* cs, const double* h, const double* shear,
                       const int32_t* type, const int32_t* state, const int32_t* boundary,
                       bool use_verlet, bool force_general, Tuning tune, HostRun& hr,
                       std::string& why) {
    if (!x || !mass || !rho || !h || !type || n_real < 0 || n_dummy < 0) {
        why = "x, mass, rho, h and type are required";
        return SPHGPU_E_ARG;
    }
    const int64_t n = n_real + n_dummy;
    if (n == 0) { why = "no particles"; return SPHGPU_E_ARG; }
    if (n >= (1LL << 31) - 1) { why = "particle count exceeds int32 range"; return SPHGPU_E_CAPACITY; }
    const int dim = cfg.dim;
    hr.n = n;
    hr.n_real = n_real;
    hr.n_dummy = n_dummy;
    Params& P = hr.P;
    fill_params_cfg(P, cfg, n, n_real);
    {
        // solid and |type| == 100 particles; pair forces between particles of
        // two positive types need at least two of them, the fluid-fluid
        // penalty two fluid (< 100) types. Each range of particles reports the
        // first positive (fluid) type it meets and whether another follows.
        struct TypeScan {
            bool solid = false, t100 = false, multi = false, multi_fluid = false;
            int first_pos = 0, first_fluid = 0;
        };
        std::vector<TypeScan> ts(16);
        const int64_t nr = host_ranges(n, [&](int64_t r, int64_t lo, int64_t hi) {
            TypeScan a;
            for (int64_t i = lo; i < hi; ++i) {
                const int t = type[i];
                a.solid |= t > 100 || t < -100;
                a.t100 |= t == 100 || t == -100;
                if (t > 0) {
                    if (a.first_pos == 0) a.first_pos = t;
                    else if (t != a.first_pos) a.multi = true;
                    if (t < 100) {
                        if (a.first_fluid == 0) a.first_fluid = t;
                        else if (t != a.first_fluid) a.multi_fluid = true;
                    }
                }
            }
            ts[r] = a;
        });
        TypeScan all;
        for (int64_t r = 0; r < nr; ++r) {
            const TypeScan& b = ts[r];
            all.solid |= b.solid;
            all.t100 |= b.t100;
            all.multi |= b.multi;
            all.multi_fluid |= b.multi_fluid;
            if (b.first_pos) {
                if (!all.first_pos) all.first_pos = b.first_pos;
                else if (b.first_pos != all.first_pos) all.multi = true;
            }
            if (b.first_fluid) {
                if (!all.first_fluid) all.first_fluid = b.first_fluid;
                else if (b.first_fluid != all.first_fluid) all.multi_fluid = true;
            }
        }
        P.has_solid = all.solid ? 1 : 0;
        P.has_t100 = all.t100 ? 1 : 0;
        P.pair_ext = cfg.external_force && (cfg.lj_r0 > 0 || all.multi);
        P.ext_pow = cfg.external_force && (cfg.lj_r0 > 0 || all.multi_fluid);
    }
    hr.fast = fast_path_ok(cfg, P, force_general);
    if (P.mixed && !hr.fast) {
        // outside the fluid set, mixed precision keeps the FP64 state of the
        // general path and runs its pair sums in FP32
        P.mixed = 0;
        P.pair32 = 1;
    }
    if (cfg.kernel_correction)
        for (int64_t i = 0; i < n_real; ++i)
            if (type[i] <= 100) {
                why = "kernel gradient correction with non-solid real particles "
                      "(the reference divides by a zero matrix there)";
                return SPHGPU_E_BAD_CONFIG;
            }
    {
        std::vector<char> bad(16, 0);
        host_ranges(n * dim, [&](int64_t r, int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; ++i)
                if (!std::isfinite(x[i])) { bad[r] = 1; break; }
        });
        for (char b : bad)
            if (b) { why = "non-finite position"; return SPHGPU_E_NONFINITE; }
    }
    double h_next_max = 0.0;
    if (P.sle == 1)   // hsml.f90:22-25, with the uploaded density
        for (int64_t i = 0; i < n_real; ++i)
            if (rho[i] > 0) h_next_max = std::max(h_next_max, 2.0 * std::pow(mass[i] / rho[i], 1.0 / dim));
    if (!(tune.skin_frac > 0)) tune.skin_frac = hr.fast ? kSkinFast : kSkinGeneral;
    plan_grid_params(P, dim, use_verlet, n, x, h, tune, h_next_max);
    // Slot order: the grid's cells in tiles of te[0] x te[1] (x te[2]) cells,
    // tiles in lexicographic order and cells lexicographic within a tile,
    // ascending particle index within a cell, by a stable counting sort over
    // the tiled cell keys. A wave, a block and an XCD's range of blocks then
    // cover compact regions, whose neighbours' records stay in that XCD's
    // L2; lexicographic order (one tile) spreads a particle's neighbours over
    // several planes of cells.
    const int T = tune.tile > 0 ? tune.tile : dim == 3 ? (hr.fast ? kTileFast3 : kTileGeneral3) : kTileWhole;
    const int TX = tune.tile_x > 0 ? tune.tile_x : (hr.fast || dim == 2) ? kTileWhole : T;
    int64_t te[3] = {1, 1, 1}, tn[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) {
        te[d] = std::min<int64_t>(d == 0 ? TX : T, P.nc[d]);
        tn[d] = (P.nc[d] + te[d] - 1) / te[d];
    }
    int64_t nkeys = tn[0] * tn[1] * tn[2] * te[0] * te[1] * te[2];
    if (nkeys > 2 * P.ncell + 4096 || nkeys >= (1LL << 31)) {
        // tiles that round the grid up this much: one tile (lexicographic)
        for (int d = 0; d < 3; ++d) { te[d] = P.nc[d]; tn[d] = 1; }
        nkeys = P.ncell;
    }
    const int64_t tcells = te[0] * te[1] * te[2];
    std::vector<uint32_t>& key = hr.sort_key;
    std::vector<uint32_t>& start = hr.sort_start;
    std::vector<int64_t>& order = hr.order;
    key.resize(n);
    // A stable counting sort in O(n) on the host threads. The key range is
    // split among O owners. Each range of particle indices computes its keys
    // and lists its indices by owner (in index order); each owner then
    // counts, scans and places its keys' particles, taking the ranges in
    // index order, so a cell keeps ascending indices.
    const int64_t O = n < (1 << 16) ? 1 : std::min<int64_t>(16, std::max(1u, std::thread::hardware_concurrency()));
    const int64_t kchunk = (nkeys + O - 1) / O;
    std::vector<std::vector<int32_t>> byown(16);
    std::vector<int64_t> ownoff(16 * 17, 0);
    const int64_t R = host_ranges(n, [&](int64_t r, int64_t lo, int64_t hi) {
        int64_t* off = &ownoff[r * 17];
        for (int64_t i = lo; i < hi; ++i) {
            int64_t cc[3] = {0, 0, 0};
            for (int d = 0; d < dim; ++d) {
                double t = std::floor((x[i * dim + d] - P.org[d]) * P.inv_cs);
                t = std::min(std::max(t, 0.0), (double)(P.nc[d] - 1));
                cc[d] = (int64_t)t;
            }
            const int64_t tile = (cc[2] / te[2] * tn[1] + cc[1] / te[1]) * tn[0] + cc[0] / te[0];
            const int64_t local = (cc[2] % te[2] * te[1] + cc[1] % te[1]) * te[0] + cc[0] % te[0];
            key[i] = (uint32_t)(tile * tcells + local);
            ++off[key[i] / kchunk + 1];
        }
        for (int64_t o = 0; o < O; ++o) off[o + 1] += off[o];
        std::vector<int32_t>& list = byown[r];
        list.resize(hi - lo);
        std::vector<int64_t> pos(off, off + O);
        for (int64_t i = lo; i < hi; ++i) list[pos[key[i] / kchunk]++] = (int32_t)i;
    });
    start.assign((size_t)nkeys + 1, 0u);
    order.resize(n);
    std::vector<int64_t> own_total(16, 0);
    auto each_own = [&](auto&& f) {
        host_tasks(O, [&](int64_t o) {
            for (int64_t r = 0; r < R; ++r) {
                const int64_t* off = &ownoff[r * 17];
                const std::vector<int32_t>& list = byown[r];
                for (int64_t q = off[o]; q < off[o + 1]; ++q) f(list[q]);
            }
        });
    };
    // counts, then each owner's exclusive scan of its keys
    each_own([&](int32_t i) { ++start[key[i]]; });
    host_tasks(O, [&](int64_t o) {
        const int64_t klo = o * kchunk, khi = std::min(nkeys, klo + kchunk);
        int64_t acc = 0;
        for (int64_t k = klo; k < khi; ++k) {
            const int64_t c = start[k];
            start[k] = (uint32_t)acc;
            acc += c;
        }
        own_total[o] = acc;
    });
    std::vector<int64_t> base(17, 0);
    for (int64_t o = 0; o < O; ++o) base[o + 1] = base[o] + own_total[o];
    host_tasks(O, [&](int64_t o) {
        const int64_t klo = o * kchunk, khi = std::min(nkeys, klo + kchunk);
        for (int64_t k = klo; k < khi; ++k) start[k] += (uint32_t)base[o];
    });
    each_own([&](int32_t i) { order[start[key[i]]++] = i; });
    hr.type_idx.assign(type, type + n);
    hr.posh.resize(n); hr.vm.resize(n); hr.rho.resize(n); hr.e.resize(n); hr.c.resize(n);
    hr.p.resize(n); hr.meta.resize(n);
    if (shear) hr.shear.resize((size_t)n * dim * dim);
    else hr.shear.clear();   // the device copy starts at zero
    host_parallel(n, [&](int64_t lo, int64_t hi) {
        for (int64_t s = lo; s < hi; ++s) {
            int64_t i = order[s];
            hr.posh[s] = make_double4(x[i * dim], x[i * dim + 1], dim == 3 ? x[i * dim + 2] : 0.0, h[i]);
            double vx = v ? v[i * dim] : 0, vy = v ? v[i * dim + 1] : 0, vz = (v && dim == 3) ? v[i * dim + 2] : 0;
            hr.vm[s] = make_double4(vx, vy, vz, mass[i]);
            hr.rho[s] = rho[i];
            hr.e[s] = e ? e[i] : 0.0;
            hr.c[s] = cs ? cs[i] : 0.0;
            hr.p[s] = p ? p[i] : 0.0;
            hr.meta[s] = make_int4(type[i], state ? state[i] : 0, boundary ? boundary[i] : 0, (int)i);
            if (shear)
                for (int q = 0; q < dim * dim; ++q) hr.shear[s * dim * dim + q] = shear[i * dim * dim + q];
        }
    });
    // slots of the real solid targets (the solid launch of the general force
    // pass covers these alone; slots do not move after upload)
    hr.sidx.clear();
    if (P.has_solid)
        for (int64_t s = 0; s < n; ++s) {
            const int4 mt = hr.meta[s];
            if (mt.w < n_real && (mt.x > 100 || mt.x < -100)) hr.sidx.push_back((int)s);
        }
    P.n_sidx = (int)hr.sidx.size();
    if (P.mixed) {
        hr.drho.resize(n);
        for (int64_t s = 0; s < n; ++s) {
            const sphgpu_material* m = nullptr;
            int a = std::abs(hr.meta[s].x);
            for (int k = 0; k < cfg.n_materials; ++k)
                if (cfg.mat[k].type == a) m = &cfg.mat[k];
            double r0 = m ? m->rho0 : 0.0;   // the same reference density eos_mixed uses
            hr.drho[s] = (float)(hr.rho[s] - r0);
        }
    }
    return 0;
}

extern "C" int sphgpu_upload(sphgpu_ctx* c, int64_t n_real, int64_t n_dummy, const double* x,
                             const double* v, const double* mass, const double* rho,
                             const double* p, const double* e, const double* cs,
                             const double* h, const double* shear, const int32_t* type,
                             const int32_t* state, const int32_t* boundary) {
    if (!c) return SPHGPU_E_ARG;
    if (n_real + n_dummy > c->capacity)
        return fail(c, SPHGPU_E_CAPACITY, "more particles than the context capacity");
    HostRun& hr = c->up;
    c->up_meta = false;
    std::string why;
    int rc = prepare_run(c->cfg, n_real, n_dummy, x, v, mass, rho, p, e, cs, h, shear, type, state,
                         boundary, c->use_verlet, c->force_general, env_tuning(), hr, why);
    if (rc) return fail(c, rc, why);
    const int64_t n = hr.n;
    const int dim = c->dim;
    c->n = n;
    c->n_real = n_real;
    c->n_dummy = n_dummy;
    c->P = hr.P;
    rc = alloc_state(c);
    if (rc) return rc;
    rc = alloc_cells(c);
    if (rc) return rc;
    // every state buffer starts at zero (alloc_state), so the displacement
    // and the derivatives need no copy
    const auto& posh = hr.posh;
    const auto& vm = hr.vm;
    const auto& vrho = hr.rho;
    const auto& ve = hr.e;
    const auto& vc = hr.c;
    const auto& vp = hr.p;
    const auto& meta = hr.meta;
    const auto& vshear = hr.shear;
    Dev& D = c->D;
    HIPCHK(c, hipMemcpyAsync(D.posh, posh.data(), n * sizeof(double4), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(D.vm, vm.data(), n * sizeof(double4), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(D.rho, vrho.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(D.e, ve.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    std::vector<float> vef;
    if (c->P.mixed) {
        vef.assign(ve.begin(), ve.end());
        HIPCHK(c, hipMemcpyAsync(D.ef, vef.data(), n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipMemcpyAsync(D.c, vc.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(D.p, vp.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(D.meta, meta.data(), n * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    if (!vshear.empty())
        HIPCHK(c, hipMemcpyAsync(D.shear, vshear.data(), vshear.size() * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
    if (!hr.sidx.empty())
        HIPCHK(c, hipMemcpyAsync(D.sidx, hr.sidx.data(), hr.sidx.size() * sizeof(int), hipMemcpyHostToDevice,
                                 c->stream));
    if (c->P.mixed) {
        HIPCHK(c, hipMemcpyAsync(D.drho, hr.drho.data(), n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    HIPCHK(c, hipMemsetAsync(D.err, 0, sizeof(int) * 4, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->step = 1;
    c->has_state = true;
    c->fused_done = false;
    drop_graph(c);
    rc = dim == 2 ? size_neighbour_list<2>(c) : size_neighbour_list<3>(c);
    if (rc) { c->has_state = false; return rc; }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_spare(c);
    c->up_meta = true;
    // the grid partials of the uploaded positions seed the first step
    const int nb = (int)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_bbox<false>, dim3(nb), dim3(kBlock), 0, c->stream, c->P, c->D, RunMap{});
    HIPCHK(c, hipGetLastError());
    return check_err(c);
}

extern "C" int sphgpu_derivatives(sphgpu_ctx* c) {
    if (!c || !c->has_state) return fail(c, SPHGPU_E_ARG, "no particle state uploaded");
    const int nb = (int)((c->P.n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_bbox<false>, dim3(nb), dim3(kBlock), 0, c->stream, c->P, c->D, RunMap{});
    // the candidate list of the fast path survives between calls (the
    // displacement vote reopens it), so the result does not depend on how
    // the host splits its calls
    if (!c->fused_done) force_rebuild(c);
    if (c->dim == 2) launch_derivs<2>(c, 2);
    else launch_derivs<3>(c, 2);
    c->fused_done = c->use_verlet;
    HIPCHK(c, hipGetLastError());
    return check_err(c);
}

extern "C" int sphgpu_advance(sphgpu_ctx* c, int32_t nsteps) {
    if (!c || !c->has_state) return fail(c, SPHGPU_E_ARG, "no particle state uploaded");
    if (nsteps < 0) return fail(c, SPHGPU_E_ARG, "negative step count");
    const int nb = (int)((c->P.n + kBlock - 1) / kBlock);
    // refresh the grid partials; build the candidate list if none is valid
    hipLaunchKernelGGL(k_bbox<false>, dim3(nb), dim3(kBlock), 0, c->stream, c->P, c->D, RunMap{});
    if (!c->fused_done) force_rebuild(c);
    if (nsteps > 0) c->fused_done = c->use_verlet;
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    double force_ms = 0;
    int64_t force_n = 0;
    if (fluid_fast_path(c) && c->use_verlet && nsteps > 0) {
        int rc = c->dim == 2 ? advance_fast<2>(c, nsteps) : advance_fast<3>(c, nsteps);
        if (rc) return rc;
    } else {
        int rc = c->dim == 2 ? advance_general<2>(c, nsteps) : advance_general<3>(c, nsteps);
        if (rc) return rc;
    }
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHK(c, hipGetLastError());
    int rc = check_err(c);
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    c->last_advance_ms = ms;
    if (nsteps > 0) {
        float fm = 0;
        hipEventElapsedTime(&fm, c->ev[2], c->ev[3]);
        force_ms = fm;
        force_n = 1;
    }
    c->last_force_ms = force_ms;
    c->last_force_launches = force_n;
    return rc;
}

// Copy one field of n particles (slots 0..n-1 of the view D) to dst in the
// caller's particle order (meta.w).
#define VCHK(call)                                                           \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) {                                              \
            msg = std::string(#call) + ": " + hipGetErrorString(e_);         \
            return SPHGPU_E_HIP;                                             \
        }                                                                    \
    } while (0)

// A slot-ordered field in particle order: dst[meta[s].w] = src[s] for
// elements of `words` 32-bit words (a run's meta.w is a permutation of 0..n-1).
// `bad` is set when a slot's particle index is out of range: the slot map
// is written once at upload from a host permutation, so this means the
// device state was overwritten, and the caller reports it instead of
// returning an incomplete field.
__global__ __launch_bounds__(kBlock) void k_particle_order(const uint32_t* __restrict__ src,
                                                           const int4* __restrict__ meta, int64_t n,
                                                           int words, uint32_t* __restrict__ dst,
                                                           int* bad) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const int64_t i = meta[s].w;
    if (i < 0 || i >= n) {
        atomicOr(bad, 1);
        return;
    }
    for (int q = 0; q < words; ++q) dst[i * words + q] = src[s * words + q];
}

// The inverse for a divergence field: the value of particle meta[s].w (host
// order, staged on the device) into slot s, as component `comp` (0: position,
// 1: velocity) of the FP64 record and, when present, of the FP32 one.
__global__ __launch_bounds__(kBlock) void k_slot_div(const double* __restrict__ src,
                                                     const int4* __restrict__ meta, int64_t n, int comp,
                                                     double2* div, float2* divf, int* bad) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const int64_t i = meta[s].w;
    if (i < 0 || i >= n) {
        atomicOr(bad, 1);
        return;
    }
    const double v = src[i];
    if (div) {
        if (comp) div[s].y = v;
        else div[s].x = v;
    }
    if (divf) {
        if (comp) divf[s].y = (float)v;
        else divf[s].x = (float)v;
    }
}

// One field of a run in particle order. `order`: device scratch of n
// elements of the widest field (n * dim * dim doubles); with it the field is
// put in particle order on the device and `type_by_index` gives each
// particle's type (for the mixed density), else the host permutes with the
// run's meta read from the device. `stage`: host staging for the copy, or
// null for a pageable buffer.
static int download_view(const Dev& D, int64_t n, int dim, bool mixed, const sphgpu_config& cfg,
                         hipStream_t stream, int32_t field, double* dst, std::string& msg,
                         void* order = nullptr, const int32_t* type_by_index = nullptr,
                         HostStage* stage = nullptr, int* bad = nullptr) {
    const bool ordered = order != nullptr && type_by_index != nullptr && bad != nullptr;
    std::vector<int4> meta_dev;
    const int4* meta = nullptr;
    if (!ordered) {
        meta_dev.resize(n);
        VCHK(hipMemcpyAsync(meta_dev.data(), D.meta, n * sizeof(int4), hipMemcpyDeviceToHost, stream));
        meta = meta_dev.data();
    }
    auto idx = [&](int64_t s) -> int64_t { return ordered ? s : (int64_t)meta[s].w; };
    auto typ = [&](int64_t s) -> int { return ordered ? type_by_index[s] : meta[s].x; };
    // copy n elements of `elem` bytes into h, in particle order when ordered
    auto fetch = [&](void* h, const void* src, size_t elem) -> int {
        if (!src) {
            msg = std::string("field not held by this run: ") + kFieldNames[field];
            return SPHGPU_E_ARG;
        }
        int flag = 0;
        if (ordered) {
            VCHK(hipMemsetAsync(bad, 0, sizeof(int), stream));
            hipLaunchKernelGGL(k_particle_order, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                               stream, static_cast<const uint32_t*>(src), D.meta, n, (int)(elem / 4),
                               static_cast<uint32_t*>(order), bad);
            VCHK(hipGetLastError());
            src = order;
        }
        VCHK(hipMemcpyAsync(h, src, n * elem, hipMemcpyDeviceToHost, stream));
        if (ordered) VCHK(hipMemcpyAsync(&flag, bad, sizeof(int), hipMemcpyDeviceToHost, stream));
        VCHK(hipStreamSynchronize(stream));
        if (flag) {
            msg = "slot-to-particle map out of range: the device state was overwritten";
            return SPHGPU_E_HIP;
        }
        return 0;
    };
    HostStage local;
    HostStage& st = stage ? *stage : local;
    bool staged = true;
    auto buf = [&](size_t bytes) -> void* {
        void* q = st.get(bytes);
        staged = q != nullptr;
        return q;
    };
#define STAGE(T, name, count)                                  \
    T* name = static_cast<T*>(buf((count) * sizeof(T)));       \
    if (!staged) {                                              \
        msg = "host staging allocation failed";                 \
        return SPHGPU_E_HIP;                                    \
    }
    auto get4 = [&](double4* src, int which) -> int {  // which: 0 xyz, 1 w
        STAGE(double4, h, n);
        if (int rc_ = fetch(h, src, sizeof(double4))) return rc_;
        host_parallel(n, [&](int64_t lo_, int64_t hi_) { for (int64_t s = lo_; s < hi_; ++s) {
            int64_t i = idx(s);
            if (which == 0) {
                dst[i * dim] = h[s].x;
                dst[i * dim + 1] = h[s].y;
                if (dim == 3) dst[i * dim + 2] = h[s].z;
            } else {
                dst[i] = h[s].w;
            }
        } });
        return 0;
    };
    auto get1 = [&](double* src) -> int {
        STAGE(double, h, n);
        if (int rc_ = fetch(h, src, sizeof(double))) return rc_;
        host_parallel(n, [&](int64_t lo_, int64_t hi_) { for (int64_t s = lo_; s < hi_; ++s) dst[idx(s)] = h[s]; });
        return 0;
    };
    auto getT = [&](double* src) -> int {
        const int64_t nn = (int64_t)dim * dim;
        if (!src) {
            std::memset(dst, 0, sizeof(double) * n * nn);
            return 0;
        }
        STAGE(double, h, n * nn);
        if (int rc_ = fetch(h, src, nn * sizeof(double))) return rc_;
        host_parallel(n, [&](int64_t lo_, int64_t hi_) { for (int64_t s = lo_; s < hi_; ++s) for (int q = 0; q < nn; ++q) dst[idx(s) * nn + q] = h[s * nn + q]; });
        return 0;
    };
    auto getf4 = [&](float4* src, int which) -> int {  // which: 0 xyz, 1 w
        STAGE(float4, h, n);
        if (int rc_ = fetch(h, src, sizeof(float4))) return rc_;
        host_parallel(n, [&](int64_t lo_, int64_t hi_) { for (int64_t s = lo_; s < hi_; ++s) {
            int64_t i = idx(s);
            if (which == 0) {
                dst[i * dim] = h[s].x;
                dst[i * dim + 1] = h[s].y;
                if (dim == 3) dst[i * dim + 2] = h[s].z;
            } else {
                dst[i] = h[s].w;
            }
        } });
        return 0;
    };
    auto getf1 = [&](float* src, bool add_rho0) -> int {
        STAGE(float, h, n);
        if (int rc_ = fetch(h, src, sizeof(float))) return rc_;
        host_parallel(n, [&](int64_t lo_, int64_t hi_) { for (int64_t s = lo_; s < hi_; ++s) {
            double r0 = 0.0;
            if (add_rho0) {
                int a = std::abs(typ(s));
                for (int k = 0; k < cfg.n_materials; ++k)
                    if (cfg.mat[k].type == a) r0 = cfg.mat[k].rho0;
            }
            dst[idx(s)] = r0 + (double)h[s];
        } });
        return 0;
    };
    if (mixed) {
        switch (field) {
        case SPHGPU_F_DVDT: return getf4(D.dvf, 0);
        case SPHGPU_F_DRHODT: return getf4(D.dvf, 1);
        case SPHGPU_F_RHO: return getf1(D.drho, true);
        case SPHGPU_F_DISP: return getf4(D.dispf, 0);
        case SPHGPU_F_E: return getf1(D.ef, false);
        case SPHGPU_F_DIVX:
        case SPHGPU_F_DIVV: {
            STAGE(float2, h, n);
            if (int rc_ = fetch(h, D.divf, sizeof(float2))) return rc_;
            host_parallel(n, [&](int64_t lo_, int64_t hi_) { for (int64_t s = lo_; s < hi_; ++s) dst[idx(s)] = field == SPHGPU_F_DIVX ? h[s].x : h[s].y; });
            return 0;
        }
        case SPHGPU_F_DEDT: return getf1(D.def, false);
        default: break;
        }
    }
    auto get2 = [&](int which) -> int {
        STAGE(double2, h, n);
        if (int rc_ = fetch(h, D.div, sizeof(double2))) return rc_;
        host_parallel(n, [&](int64_t lo_, int64_t hi_) { for (int64_t s = lo_; s < hi_; ++s) dst[idx(s)] = which ? h[s].y : h[s].x; });
        return 0;
    };
    switch (field) {
    case SPHGPU_F_X: return get4(D.posh, 0);
    case SPHGPU_F_H: return get4(D.posh, 1);
    case SPHGPU_F_V: return get4(D.vm, 0);
    case SPHGPU_F_MASS: return get4(D.vm, 1);
    case SPHGPU_F_DISP: return get4(D.disp, 0);
    case SPHGPU_F_DVDT: return get4(D.dv, 0);
    case SPHGPU_F_DRHODT: return get4(D.dv, 1);
    case SPHGPU_F_RHO: return get1(D.rho);
    case SPHGPU_F_P: return get1(D.p);
    case SPHGPU_F_E: return get1(D.e);
    case SPHGPU_F_C: return get1(D.c);
    case SPHGPU_F_DEDT: return get1(D.de);
    case SPHGPU_F_ETA: return get1(D.eta);
    case SPHGPU_F_STRESS: return getT(D.stress);
    case SPHGPU_F_SHEAR: return getT(D.shear);
    case SPHGPU_F_DSDT: return getT(D.dsdt);
    case SPHGPU_F_DIVX: return get2(0);
    case SPHGPU_F_DIVV: return get2(1);
    default: break;
    }
#undef STAGE
    msg = std::string("field not available: ") + kFieldNames[field];
    return SPHGPU_E_ARG;
}


extern "C" int sphgpu_download(sphgpu_ctx* c, int32_t field, double* dst) {
    if (!c || !c->has_state) return fail(c, SPHGPU_E_ARG, "no particle state uploaded");
    if (!dst || field < 0 || field >= SPHGPU_F_NFIELDS) return fail(c, SPHGPU_E_ARG, "bad field");
    std::string msg;
    // device scratch for the reorder, the size of the widest field at capacity
    const size_t obytes = (size_t)c->capacity * c->dim * c->dim * sizeof(double);
    if (c->up_meta && !c->dorder) HIPCHK(c, hipMalloc(&c->dorder, obytes));
    const bool typed = c->up_meta && (int64_t)c->up.type_idx.size() == c->n;
    int rc = download_view(c->D, c->n, c->dim, c->P.mixed != 0, c->cfg, c->stream, field, dst, msg,
                           typed ? c->dorder : nullptr, typed ? c->up.type_idx.data() : nullptr, &c->dstage,
                           c->D.err + 3);
    return rc ? fail(c, rc, msg) : 0;
}

// One field of the resident state from host memory in upload order. Only
// the divergences, which a caller that uploads on every step needs to carry
// from the previous step (Shepard normalisation reads the last step's
// divergencePosition, density.f90:25-37).
extern "C" int sphgpu_upload_field(sphgpu_ctx* c, int32_t field, const double* src) {
    if (!c || !c->has_state) return fail(c, SPHGPU_E_ARG, "no particle state uploaded");
    if (!src) return fail(c, SPHGPU_E_ARG, "NULL source");
    if (field != SPHGPU_F_DIVX && field != SPHGPU_F_DIVV)
        return fail(c, SPHGPU_E_ARG, std::string("field cannot be uploaded alone: ") +
                                         (field >= 0 && field < SPHGPU_F_NFIELDS ? kFieldNames[field] : "?"));
    const size_t obytes = (size_t)c->capacity * c->dim * c->dim * sizeof(double);
    if (!c->dorder) HIPCHK(c, hipMalloc(&c->dorder, obytes));
    int* bad = c->D.err + 3;
    int flag = 0;
    HIPCHK(c, hipMemsetAsync(bad, 0, sizeof(int), c->stream));
    HIPCHK(c, hipMemcpyAsync(c->dorder, src, c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_slot_div, dim3((unsigned)((c->n + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->stream,
                       static_cast<const double*>(c->dorder), c->D.meta, c->n, field == SPHGPU_F_DIVV ? 1 : 0,
                       c->D.div, c->D.divf, bad);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(&flag, bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (flag) return fail(c, SPHGPU_E_HIP, "slot-to-particle map out of range: the device state was overwritten");
    return 0;
}

extern "C" int sphgpu_export_pairs(sphgpu_ctx* c, int64_t* offsets, int32_t* nbr, int64_t cap,
                                   int64_t* needed) {
    if (!c || !c->has_state) return fail(c, SPHGPU_E_ARG, "no particle state uploaded");
    if (!offsets) return fail(c, SPHGPU_E_ARG, "offsets required");
    const int nb = (int)((c->P.n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_bbox<false>, dim3(nb), dim3(kBlock), 0, c->stream, c->P, c->D, RunMap{});
    if (c->dim == 2) {
        hipLaunchKernelGGL((k_grid<2, false>), dim3(1), dim3(kGridThreads), 0, c->stream, c->P, c->D, RunMap{});
        launch_sort<2>(c);
        launch_point_build<2>(c, 2);
    } else {
        hipLaunchKernelGGL((k_grid<3, false>), dim3(1), dim3(kGridThreads), 0, c->stream, c->P, c->D, RunMap{});
        launch_sort<3>(c);
        launch_point_build<3>(c, 2);
    }
    HIPCHK(c, hipGetLastError());
    int rc = check_err(c);
    if (rc) return rc;
    const int64_t n = c->n, nr = c->n_real;
    std::vector<int4> meta(n);
    std::vector<int> cnt(n);
    HIPCHK(c, hipMemcpyAsync(meta.data(), c->D.meta, n * sizeof(int4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(cnt.data(), c->D.ncnt, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<int64_t> slot_of(nr, -1);
    for (int64_t s = 0; s < n; ++s)
        if (meta[s].w < nr) slot_of[meta[s].w] = s;
    offsets[0] = 0;
    for (int64_t i = 0; i < nr; ++i) {
        int64_t s = slot_of[i];
        int k = (meta[s].y == -1) ? 0 : cnt[s];
        offsets[i + 1] = offsets[i] + k;
    }
    if (needed) *needed = offsets[nr];
    if (!nbr || offsets[nr] > cap) return fail(c, SPHGPU_E_PAIR_CAPACITY, "pair buffer too small");
    const int64_t lcap = c->P.nbr_cap;
    std::vector<int> all((size_t)vlist_entries(n, (int)lcap));
    HIPCHK(c, hipMemcpyAsync(all.data(), c->D.nbr, all.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < nr; ++i) {
        int64_t s = slot_of[i];
        int64_t o = offsets[i];
        int k = (int)(offsets[i + 1] - offsets[i]);
        for (int q = 0; q < k; ++q) nbr[o + q] = meta[all[(size_t)vlist_at((int)s, q, (int)lcap)]].w;
        std::sort(nbr + o, nbr + o + k);
    }
    return 0;
}

extern "C" int sphgpu_neighbour_stats(sphgpu_ctx* c, int64_t* out4) {
    if (!c || !c->has_state || !out4) return fail(c, SPHGPU_E_ARG, "bad argument");
    unsigned long long* d = nullptr;
    unsigned long long init[4] = {0, 0, ~0ull, 0};
    HIPCHK(c, hipMalloc(&d, sizeof(init)));
    HIPCHK(c, hipMemcpyAsync(d, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
    const int nb = std::min((int)((c->P.n + kBlock - 1) / kBlock), kStatsBlocks);
    const int* counts = (fluid_fast_path(c) && c->use_verlet && c->fused_done) ? c->D.ecnt : c->D.ncnt;
    hipLaunchKernelGGL(k_stats, dim3(nb), dim3(kBlock), 0, c->stream, c->P, c->D, counts, d);
    unsigned long long h[4];
    hipError_t e1 = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream);
    hipError_t e2 = hipStreamSynchronize(c->stream);
    hipFree(d);
    if (e1 != hipSuccess || e2 != hipSuccess) return fail(c, SPHGPU_E_HIP, "neighbour statistics");
    for (int k = 0; k < 4; ++k) out4[k] = (int64_t)h[k];
    if (c->n_real == 0) out4[2] = 0;
    return 0;
}

extern "C" int sphgpu_grid_info(sphgpu_ctx* c, double* out9) {
    if (!c || !c->has_state || !out9) return fail(c, SPHGPU_E_ARG, "bad argument");
    RefGrid g;
    HIPCHK(c, hipMemcpyAsync(&g, c->D.grid, sizeof(RefGrid), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int d = 0; d < 3; ++d) {
        out9[d] = g.minc[d];
        out9[3 + d] = g.maxc[d];
        out9[6 + d] = g.num[d];
    }
    return 0;
}

extern "C" int sphgpu_timing(sphgpu_ctx* c, double* advance_ms, double* force_kernel_ms,
                             int64_t* force_launches) {
    if (!c) return SPHGPU_E_ARG;
    if (advance_ms) *advance_ms = c->last_advance_ms;
    if (force_kernel_ms) *force_kernel_ms = c->last_force_ms;
    if (force_launches) *force_launches = c->last_force_launches;
    return 0;
}

extern "C" int64_t sphgpu_step(sphgpu_ctx* c) { return c ? c->step : -1; }

extern "C" const char* sphgpu_last_error(sphgpu_ctx* c) {
    if (c) return c->err.c_str();
    return g_create_error.empty() ? "no context" : g_create_error.c_str();
}

extern "C" void sphgpu_destroy(sphgpu_ctx* c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    drop_graph(c);
    free_all(c);
    c->dstage.release();
    if (c->dorder) hipFree(c->dorder);
    for (auto& e : c->ev)
        if (e) hipEventDestroy(e);
    if (c->fork) hipEventDestroy(c->fork);
    if (c->join) hipEventDestroy(c->join);
    if (c->h_err) hipHostFree(c->h_err);
    if (c->side) {
        hipStreamSynchronize(c->side);
        hipStreamDestroy(c->side);
    }
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

#include "checkpoint.hpp"
#include "batch.hpp"
#include "noalloc.hpp"
