// noalloc.hpp - the BGGS pair export with no allocation: sphgpu_pairs_noalloc.
//
// The same result as sphgpu_export_pairs (for each real particle i the
// neighbour ids of the reference's BGGS, asymmetricSearch.f90:160-257, in
// ascending order, as CSR offsets + ids) computed from caller-owned device
// arrays, with no context and no device allocation: every intermediate
// (box partials, the reference grid, cell counts and starts, bucket items,
// scan partials, the status word) lives in one file-scope __device__ array,
// `scratch`, of kNaScratchBytes. Its device address is taken once per device
// with hipGetSymbolAddress on the first call and used from then on; the
// array's name is never passed to a kernel.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one
// device (two streams, two host threads) race and must be serialised by the
// caller. Everything a call needs to start from zero is cleared by its own
// fill kernel (no memset), so the last call's contents do not matter.
//
// Pipeline (one stream, kernels only, one synchronisation at the end for the
// status and the pair count):
//   k_na_init      status word cleared
//   k_na_box       per-block bounding-box partials over all n particles
//   k_na_grid      the reference grid (asymmetricSearch.f90:49-131)
//   k_na_zero      cell counts and fill cursors cleared
//   k_na_locate    cell of each source (State /= -1), counts per cell
//   scan           cell starts (exclusive prefix sum)
//   k_na_fill      bucket items
//   k_na_count     accepted neighbours per target, into offsets[i]
//   scan           offsets (exclusive prefix sum; offsets[n_real] = total)
//   k_na_write     the ids (when the total fits cap), sorted per target

constexpr int64_t kNaMaxN = int64_t(1) << 20;              // particles (real + boundary) per call
constexpr int64_t kNaMaxCells = kNaMaxN + (int64_t(1) << 16);
constexpr int kNaBoxBlocks = (int)((kNaMaxN + kBlock - 1) / kBlock);
constexpr int kNaScanItems = 4 * kBlock;                    // items per scan block
constexpr int64_t kNaScanParts = (kNaMaxCells + 1 + kNaScanItems - 1) / kNaScanItems + 1;

constexpr int kErrNaCells = 16;   // the reference grid has more cells than the scratch holds

struct NaGrid {
    double minc[3], maxc[3];
    int num[3];
    int status;        // kErr* bits of this call
    int64_t ncell;
};

// scratch layout (byte offsets, 256-aligned)
constexpr size_t na_align(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t kNaOffGrid = 0;
constexpr size_t kNaOffBox = na_align(sizeof(NaGrid));
constexpr size_t kNaOffCellCnt = na_align(kNaOffBox + sizeof(double) * 6 * kNaBoxBlocks);
constexpr size_t kNaOffCellFill = na_align(kNaOffCellCnt + sizeof(int) * (kNaMaxCells + 1));
constexpr size_t kNaOffItems = na_align(kNaOffCellFill + sizeof(int) * kNaMaxCells);
constexpr size_t kNaOffCellOf = na_align(kNaOffItems + sizeof(int) * kNaMaxN);
constexpr size_t kNaOffParts = na_align(kNaOffCellOf + sizeof(int) * kNaMaxN);
constexpr size_t kNaScratchBytes = na_align(kNaOffParts + sizeof(int64_t) * kNaScanParts);

static __device__ __align__(256) unsigned char scratch[kNaScratchBytes];

struct NaView {
    NaGrid* grid;
    double* box;
    int* cell_cnt;     // counts, then starts (ncell + 1)
    int* cell_fill;
    int* items;
    int* cell_of;
    int64_t* parts;
};

static NaView na_view(unsigned char* base) {
    NaView v;
    v.grid = reinterpret_cast<NaGrid*>(base + kNaOffGrid);
    v.box = reinterpret_cast<double*>(base + kNaOffBox);
    v.cell_cnt = reinterpret_cast<int*>(base + kNaOffCellCnt);
    v.cell_fill = reinterpret_cast<int*>(base + kNaOffCellFill);
    v.items = reinterpret_cast<int*>(base + kNaOffItems);
    v.cell_of = reinterpret_cast<int*>(base + kNaOffCellOf);
    v.parts = reinterpret_cast<int64_t*>(base + kNaOffParts);
    return v;
}

__global__ void k_na_init(NaView v) {
    if (threadIdx.x == 0) {
        v.grid->status = 0;
        v.grid->ncell = 0;
    }
}

__global__ __launch_bounds__(kBlock) void k_na_box(NaView v, const double* __restrict__ x, int n, int dim) {
    __shared__ double smn[3][kBlock], smx[3][kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (i < n)
        for (int d = 0; d < dim; ++d) {
            const double q = x[(int64_t)i * dim + d];
            if (!isfinite(q)) atomicOr(&v.grid->status, kErrNonFinite);
            mn[d] = q;
            mx[d] = q;
        }
    for (int d = 0; d < 3; ++d) { smn[d][threadIdx.x] = mn[d]; smx[d][threadIdx.x] = mx[d]; }
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int d = 0; d < 3; ++d) {
                smn[d][threadIdx.x] = fmin(smn[d][threadIdx.x], smn[d][threadIdx.x + o]);
                smx[d][threadIdx.x] = fmax(smx[d][threadIdx.x], smx[d][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) {
            v.box[blockIdx.x * 6 + d] = smn[d][0];
            v.box[blockIdx.x * 6 + 3 + d] = smx[d][0];
        }
}

// asymmetricSearch.f90:49-131, as k_grid computes it, from nb box partials.
__global__ __launch_bounds__(kBlock) void k_na_grid(NaView v, int n, int nb, int dim) {
    __shared__ double smn[3][kBlock], smx[3][kBlock];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nb; b += kBlock)
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin(mn[d], v.box[b * 6 + d]);
            mx[d] = fmax(mx[d], v.box[b * 6 + 3 + d]);
        }
    for (int d = 0; d < 3; ++d) { smn[d][threadIdx.x] = mn[d]; smx[d][threadIdx.x] = mx[d]; }
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int d = 0; d < 3; ++d) {
                smn[d][threadIdx.x] = fmin(smn[d][threadIdx.x], smn[d][threadIdx.x + o]);
                smx[d][threadIdx.x] = fmax(smx[d][threadIdx.x], smx[d][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    NaGrid g = *v.grid;
    double delta[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < 3; ++d) { g.minc[d] = 0.0; g.maxc[d] = 0.0; g.num[d] = 1; }
    for (int d = 0; d < dim; ++d) {
        delta[d] = rn_sub(smx[d][0], smn[d][0]);
        g.maxc[d] = rn_add(smx[d][0], rn_mul(kGmrM1, delta[d]));
        g.minc[d] = rn_sub(smn[d][0], rn_mul(kGmrM1, delta[d]));
    }
    const double N = (double)n;
    if (dim == 2) {
        double t = __dsqrt_rn(__ddiv_rn(rn_mul(N, delta[0]), rn_mul(delta[1], (double)kNpg)));
        int c1 = (t >= 0.0 && t < 1e9) ? (int)t + 1 : 0;
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
    // a flat bounding box divides by zero in the cell counts (undefined in
    // the reference): reported, and the grid kept in bounds
    for (int d = 0; d < dim; ++d)
        if (!(delta[d] > 0.0) || g.num[d] < 1) {
            g.status |= kErrOutOfRange;
            g.num[d] = 1;
            g.maxc[d] = g.minc[d] + 1.0;
        }
    g.ncell = (int64_t)g.num[0] * g.num[1] * g.num[2];
    if (g.ncell > kNaMaxCells) {
        g.status |= kErrNaCells;
        g.ncell = 0;
    }
    *v.grid = g;
}

__global__ __launch_bounds__(kBlock) void k_na_zero(NaView v) {
    if (v.grid->status) return;
    const int64_t m = v.grid->ncell + 1;
    for (int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x; c < m; c += (int64_t)gridDim.x * kBlock) {
        v.cell_cnt[c] = 0;
        if (c < m - 1) v.cell_fill[c] = 0;
    }
}

// asymmetricSearch.f90:27-47. A position at the enlarged box's upper edge
// would land one cell past the grid (the box's 20 % margin keeps finite
// positions inside); it is reported like any position outside.
__device__ __forceinline__ bool na_locate(const NaGrid& g, const double* xi, int dim, int c[3]) {
    bool bad = false;
    c[0] = c[1] = c[2] = 1;
    for (int d = 0; d < dim; ++d) {
        c[d] = ref_locate(xi[d], g.minc[d], g.maxc[d], g.num[d], &bad);
        if (c[d] < 1 || c[d] > g.num[d]) bad = true;
    }
    return !bad;
}

__device__ __forceinline__ int64_t na_lin(const NaGrid& g, int cx, int cy, int cz) {
    return (int64_t)(cx - 1) + (int64_t)g.num[0] * ((int64_t)(cy - 1) + (int64_t)g.num[1] * (cz - 1));
}

__global__ __launch_bounds__(kBlock) void k_na_locate(NaView v, const double* __restrict__ x,
                                                      const int* __restrict__ state, int n, int dim) {
    if (v.grid->status) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    v.cell_of[i] = -1;
    if (state && state[i] == -1) return;
    const NaGrid g = *v.grid;
    double xi[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < dim; ++d) xi[d] = x[(int64_t)i * dim + d];
    int c[3];
    if (!na_locate(g, xi, dim, c)) {
        atomicOr(&v.grid->status, kErrOutOfRange);
        return;
    }
    const int64_t lin = na_lin(g, c[0], c[1], c[2]);
    v.cell_of[i] = (int)lin;
    atomicAdd(&v.cell_cnt[lin], 1);
}

// Exclusive prefix sum of len entries (len from the device: *len_dev, or
// len_host when len_dev is null), in three launches over fixed grids.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_na_scan_blocks(NaView v, T* data, const int64_t* len_dev,
                                                           int64_t len_host, int64_t* parts) {
    if (v.grid->status) return;
    const int64_t len = len_dev ? *len_dev + 1 : len_host;
    __shared__ T s[kBlock];
    const int64_t b0 = (int64_t)blockIdx.x * kNaScanItems;
    T loc[4];
    T sum = 0;
    for (int q = 0; q < 4; ++q) {
        const int64_t k = b0 + (int64_t)threadIdx.x * 4 + q;
        loc[q] = k < len ? data[k] : T(0);
        sum += loc[q];
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 1; o < kBlock; o <<= 1) {   // inclusive scan of the thread sums
        T t = threadIdx.x >= o ? s[threadIdx.x - o] : T(0);
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    T run = s[threadIdx.x] - sum;
    for (int q = 0; q < 4; ++q) {
        const int64_t k = b0 + (int64_t)threadIdx.x * 4 + q;
        if (k < len) data[k] = run;
        run += loc[q];
    }
    if (threadIdx.x == kBlock - 1) parts[blockIdx.x] = (int64_t)s[kBlock - 1];
}

__global__ __launch_bounds__(kBlock) void k_na_scan_parts(NaView v, int64_t* parts, int nparts) {
    if (v.grid->status) return;
    if (threadIdx.x != 0) return;
    int64_t run = 0;
    for (int b = 0; b < nparts; ++b) {
        const int64_t t = parts[b];
        parts[b] = run;
        run += t;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_na_scan_add(NaView v, T* data, const int64_t* len_dev,
                                                        int64_t len_host, const int64_t* parts) {
    if (v.grid->status) return;
    const int64_t len = len_dev ? *len_dev + 1 : len_host;
    const int64_t b0 = (int64_t)blockIdx.x * kNaScanItems;
    const T add = (T)parts[blockIdx.x];
    for (int q = 0; q < 4; ++q) {
        const int64_t k = b0 + (int64_t)threadIdx.x * 4 + q;
        if (k < len) data[k] += add;
    }
}

__global__ __launch_bounds__(kBlock) void k_na_fill(NaView v, int n) {
    if (v.grid->status) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int c = v.cell_of[i];
    if (c < 0) return;
    v.items[v.cell_cnt[c] + atomicAdd(&v.cell_fill[c], 1)] = i;
}

// The reference's acceptance for target i (asymmetricSearch.f90:202-248):
// window of int(h_i / span * num) + 1 cells about i's cell, clipped; skip i
// itself (State == -1 sources are not in the buckets); dx = x_i - x_j, the
// squared distance summed in axis order, correctly rounded sqrt,
// r < scale_k h_i. WRITE: store accepted ids from out[0] on.
template <bool WRITE>
__device__ __forceinline__ int na_neighbours(const NaView& v, const NaGrid& g, const double* __restrict__ x,
                                             const double* __restrict__ h, int i, int dim, int scale_k,
                                             int* out) {
    double xi[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < dim; ++d) xi[d] = x[(int64_t)i * dim + d];
    int c[3];
    if (!na_locate(g, xi, dim, c)) {
        atomicOr(&v.grid->status, kErrOutOfRange);
        return 0;
    }
    const double hi = h[i];
    const double cut = rn_mul(hi, (double)scale_k);
    int lo[3] = {1, 1, 1}, hi_c[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) {
        const int w = (int)rn_mul(__ddiv_rn(hi, rn_sub(g.maxc[d], g.minc[d])), (double)g.num[d]) + 1;
        hi_c[d] = min(c[d] + w, g.num[d]);
        lo[d] = max(c[d] - w, 1);
    }
    int cnt = 0;
    for (int cz = lo[2]; cz <= hi_c[2]; ++cz)
        for (int cy = lo[1]; cy <= hi_c[1]; ++cy)
            for (int cx = lo[0]; cx <= hi_c[0]; ++cx) {
                const int64_t lin = na_lin(g, cx, cy, cz);
                const int b = v.cell_cnt[lin], e = v.cell_cnt[lin + 1];
                for (int k = b; k < e; ++k) {
                    const int j = v.items[k];
                    if (j == i) continue;
                    double dx = rn_sub(xi[0], x[(int64_t)j * dim]);
                    double dr = rn_mul(dx, dx);
                    for (int d = 1; d < dim; ++d) {
                        dx = rn_sub(xi[d], x[(int64_t)j * dim + d]);
                        dr = rn_add(dr, rn_mul(dx, dx));
                    }
                    if (__dsqrt_rn(dr) < cut) {
                        if (WRITE) out[cnt] = j;
                        ++cnt;
                    }
                }
            }
    return cnt;
}

__global__ __launch_bounds__(kBlock) void k_na_count(NaView v, const double* __restrict__ x,
                                                     const double* __restrict__ h, const int* __restrict__ state,
                                                     int n_real, int dim, int scale_k, int64_t* offsets) {
    if (v.grid->status) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i > n_real) return;
    if (i == n_real || (state && state[i] == -1)) {
        offsets[i] = 0;
        return;
    }
    const NaGrid g = *v.grid;
    offsets[i] = na_neighbours<false>(v, g, x, h, i, dim, scale_k, nullptr);
}

__global__ __launch_bounds__(kBlock) void k_na_write(NaView v, const double* __restrict__ x,
                                                     const double* __restrict__ h, const int* __restrict__ state,
                                                     int n_real, int dim, int scale_k,
                                                     const int64_t* __restrict__ offsets, int* nbr, int64_t cap) {
    if (v.grid->status || !nbr || offsets[n_real] > cap) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_real || (state && state[i] == -1)) return;
    const NaGrid g = *v.grid;
    int* out = nbr + offsets[i];
    const int m = na_neighbours<true>(v, g, x, h, i, dim, scale_k, out);
    for (int a = 1; a < m; ++a) {       // ascending ids, as sphgpu_export_pairs returns them
        const int t = out[a];
        int b = a - 1;
        while (b >= 0 && out[b] > t) {
            out[b + 1] = out[b];
            --b;
        }
        out[b + 1] = t;
    }
}

constexpr int kNaMaxDevices = 64;
static std::mutex g_na_mutex;
static unsigned char* g_na_base[kNaMaxDevices] = {};

// the message goes where sphgpu_last_error(NULL) finds it
static int na_fail(int code, const std::string& msg) {
    g_create_error = "sphgpu_pairs_noalloc: " + msg;
    return code;
}

// The device address of `scratch` on the current device, looked up once per
// device (the array's name is a host-side stand-in, not a device address).
static int na_scratch_base(unsigned char** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return na_fail(SPHGPU_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kNaMaxDevices) return na_fail(SPHGPU_E_HIP, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_base[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return na_fail(SPHGPU_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_na_base[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_na_base[dev];
    return 0;
}

template <typename T>
static void na_scan(hipStream_t st, const NaView& v, T* data, const int64_t* len_dev, int64_t len_host,
                    int64_t max_len) {
    const int nparts = (int)((max_len + kNaScanItems - 1) / kNaScanItems);
    hipLaunchKernelGGL(k_na_scan_blocks<T>, dim3(nparts), dim3(kBlock), 0, st, v, data, len_dev, len_host, v.parts);
    hipLaunchKernelGGL(k_na_scan_parts, dim3(1), dim3(kBlock), 0, st, v, v.parts, nparts);
    hipLaunchKernelGGL(k_na_scan_add<T>, dim3(nparts), dim3(kBlock), 0, st, v, data, len_dev, len_host,
                       (const int64_t*)v.parts);
}

extern "C" int sphgpu_pairs_noalloc(int32_t dim, int32_t skf, int64_t n_real, int64_t n, const double* x,
                                    const double* h, const int32_t* state, int64_t* offsets, int32_t* nbr,
                                    int64_t cap, int64_t* needed, void* stream) {
    if (dim != 2 && dim != 3) return na_fail(SPHGPU_E_ARG, "dim must be 2 or 3");
    if (skf < 1 || skf > 3) return na_fail(SPHGPU_E_ARG, "skf must be 1, 2 or 3");
    if (n < 1 || n_real < 0 || n_real > n) return na_fail(SPHGPU_E_ARG, "need 0 <= n_real <= n and n >= 1");
    if (n > kNaMaxN)
        return na_fail(SPHGPU_E_CAPACITY, "more particles than the no-allocation search holds (" +
                                              std::to_string(kNaMaxN) + ")");
    if (!x || !h || !offsets || cap < 0 || (cap > 0 && !nbr)) return na_fail(SPHGPU_E_ARG, "bad pointer or capacity");
    unsigned char* base = nullptr;
    if (int rc = na_scratch_base(&base)) return rc;
    const NaView v = na_view(base);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int nn = (int)n, nr = (int)n_real;
    const int scale_k = skf == 1 ? 2 : 3;   // asymmetricSearch.f90:178-185
    const int nb = (nn + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_na_init, dim3(1), dim3(64), 0, st, v);
    hipLaunchKernelGGL(k_na_box, dim3(nb), dim3(kBlock), 0, st, v, x, nn, (int)dim);
    hipLaunchKernelGGL(k_na_grid, dim3(1), dim3(kBlock), 0, st, v, nn, nb, (int)dim);
    hipLaunchKernelGGL(k_na_zero, dim3(1024), dim3(kBlock), 0, st, v);
    hipLaunchKernelGGL(k_na_locate, dim3(nb), dim3(kBlock), 0, st, v, x, (const int*)state, nn, (int)dim);
    na_scan<int>(st, v, v.cell_cnt, &v.grid->ncell, 0, kNaMaxCells + 1);
    hipLaunchKernelGGL(k_na_fill, dim3(nb), dim3(kBlock), 0, st, v, nn);
    const int nbr_blocks = (nr + 1 + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_na_count, dim3(nbr_blocks), dim3(kBlock), 0, st, v, x, h, (const int*)state, nr,
                       (int)dim, scale_k, offsets);
    na_scan<int64_t>(st, v, offsets, nullptr, n_real + 1, n_real + 1);
    hipLaunchKernelGGL(k_na_write, dim3(nbr_blocks), dim3(kBlock), 0, st, v, x, h, (const int*)state, nr,
                       (int)dim, scale_k, (const int64_t*)offsets, (int*)nbr, cap);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return na_fail(SPHGPU_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    int status = 0;
    int64_t total = 0;
    e = hipMemcpyAsync(&status, &v.grid->status, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&total, offsets + n_real, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return na_fail(SPHGPU_E_HIP, std::string("status copy: ") + hipGetErrorString(e));
    if (status & kErrNonFinite) return na_fail(SPHGPU_E_NONFINITE, "non-finite position");
    if (status & kErrNaCells)
        return na_fail(SPHGPU_E_CAPACITY, "the reference grid has more cells than the no-allocation search holds");
    if (status & kErrOutOfRange)
        return na_fail(SPHGPU_E_OUT_OF_RANGE, "particle out of range (flat bounding box or position outside the grid)");
    if (needed) *needed = total;
    if (total > cap) return na_fail(SPHGPU_E_PAIR_CAPACITY, "pair buffer too small");
    return 0;
}

