// Batches of independent runs (included by sphgpu.hip after the single-run
// context). The reference runs one case per process (src/main.f90:16-22);
// serving many uneven runs on one GPU needs them resident side by side:
//
//  * one arena of particle slots and cell-table entries, allocated once;
//    a run gets a contiguous slot range (kBlock-aligned) and a cell range,
//    first fit, so admission is a question of free arena space;
//  * a device run table (one Params row per run: its grid, constants, slot
//    and cell offsets, and its own gate / error / grid words);
//  * every fast-path launch covers all runnable runs: per-particle kernels
//    map a block to (run, block within run), slot-strided kernels look the
//    run of a slot up, and k_grid takes one block per run;
//  * the cell sort is shared (one scan over all runs' cell ranges) but each
//    run rebuilds its candidate list only on its own displacement vote, so
//    a run's arithmetic does not depend on which other runs are resident.
//
// A run advances until its own last step and is then left in place for
// download; a run whose error word is set stops alone.

namespace {

// First-fit allocator of [0, cap) with coalescing of freed ranges.
struct RangeAlloc {
    std::vector<std::pair<int64_t, int64_t>> free_;   // (lo, hi), sorted, disjoint
    std::vector<std::pair<int64_t, int64_t>> used_;   // allocated (lo, hi)
    int64_t cap = 0;

    void init(int64_t c) {
        cap = c;
        free_.assign(1, {0, c});
        used_.clear();
    }
    int64_t alloc(int64_t n) {
        for (size_t k = 0; k < free_.size(); ++k) {
            auto& f = free_[k];
            if (f.second - f.first >= n) {
                int64_t lo = f.first;
                f.first += n;
                if (f.first == f.second) free_.erase(free_.begin() + k);
                used_.push_back({lo, lo + n});
                return lo;
            }
        }
        return -1;
    }
    void release(int64_t lo) {
        for (size_t k = 0; k < used_.size(); ++k)
            if (used_[k].first == lo) {
                std::pair<int64_t, int64_t> r = used_[k];
                used_.erase(used_.begin() + k);
                auto it = std::lower_bound(free_.begin(), free_.end(), r);
                it = free_.insert(it, r);
                size_t i = it - free_.begin();
                if (i + 1 < free_.size() && free_[i].second == free_[i + 1].first) {
                    free_[i].second = free_[i + 1].second;
                    free_.erase(free_.begin() + i + 1);
                }
                if (i > 0 && free_[i - 1].second == free_[i].first) {
                    free_[i - 1].second = free_[i].second;
                    free_.erase(free_.begin() + i);
                }
                return;
            }
    }
    int64_t high_water() const {
        int64_t h = 0;
        for (auto& u : used_) h = std::max(h, u.second);
        return h;
    }
    int64_t in_use() const {
        int64_t t = 0;
        for (auto& u : used_) t += u.second - u.first;
        return t;
    }
};

struct BatchRun {
    bool used = false;       // the row holds a resident run
    bool runnable = false;   // still advancing (not finished, not failed)
    int status = 0;          // sphgpu_status of a failure, 0 otherwise
    std::string msg;
    sphgpu_config cfg{};
    Params P{};
    int64_t n = 0, n_real = 0;
    int64_t s0 = 0, nslot = 0;
    int64_t c0 = 0, ncells = 0;
    int64_t step = 1;        // number of the next step
    int64_t target = 0;      // last step of the run
    std::vector<int32_t> type_idx;   // type by particle index (downloads)
};

constexpr int kBatchVerletCap2D = 48;
constexpr int kBatchVerletCap3D = 160;

// Device bytes of an arena (what sphgpu_batch_create allocates).
int64_t batch_bytes(int dim, int precision, int64_t slots, int64_t cells, int max_runs, int vcap) {
    const int64_t N = slots;
    int64_t b = 0;
    b += N * (sizeof(double4) * 3);                    // posh vm vprev
    b += N * (sizeof(double) * 6);                     // rho e c p eprev eta
    b += N * (sizeof(int4) + sizeof(double2));         // meta div
    b += N * (sizeof(uint32_t) * 2 + sizeof(int) * 4); // key refcell perm vcnt ecnt runof
    b += N * (sizeof(float4));                         // cpos
    b += N * (int64_t)vcap * sizeof(int);              // candidate lists
    if (precision == SPHGPU_MIXED)
        b += N * (sizeof(float4) * 6 + sizeof(float) * 5 + sizeof(float2));   // + ef eprevf divf
    else
        b += N * (sizeof(double4) * 4 + sizeof(double));     // dv th disp xb, de
    b += (N / kBlock + 1) * 6 * sizeof(double) + (N / kBlock + 1) * sizeof(int2);
    b += (cells + 2) * sizeof(int) * 2 + ((cells + kScanTile) / kScanTile + 1) * sizeof(int);
    b += (int64_t)max_runs * (sizeof(Params) + sizeof(RefGrid) + 2 * sizeof(unsigned long long) +
                              8 * sizeof(int) + SPHGPU_MAX_MATERIALS * sizeof(Material) + sizeof(int));
    (void)dim;
    return b;
}

}  // namespace

struct sphgpu_batch {
    int dim = 2, precision = 0, max_runs = 0, verlet_cap = 0;
    int64_t slot_cap = 0, cell_cap = 0;
    hipStream_t stream = nullptr;
    Dev D{};
    std::vector<void*> allocs;
    Params* d_runs = nullptr;
    int2* d_blk = nullptr;
    int* d_runof = nullptr;
    int* d_rows = nullptr;
    RefGrid* d_grid = nullptr;
    unsigned long long* d_vflag = nullptr;   // 2 per run
    unsigned long long* d_vany = nullptr;    // 2, batch-wide
    int* d_err = nullptr;                    // 4 per run
    int* d_maxcnt = nullptr;                 // 4 per run
    Material* d_mat = nullptr;               // SPHGPU_MAX_MATERIALS per run
    Params* h_runs = nullptr;                // pinned staging of the run table
    int* h_err = nullptr;                    // pinned copy of the error words
    RangeAlloc slots, cells;
    void* dorder = nullptr;      // device scratch: a downloaded field in particle order
    size_t dorder_bytes = 0;
    HostStage dstage{true};      // pinned staging of downloads
    std::vector<BatchRun> runs;
    std::vector<int2> blk;
    std::vector<int> rows;
    int nact = 0;
    bool map_dirty = true;
    hipEvent_t ev[2]{};
    double last_ms = 0;
    int64_t last_steps = 0, last_particle_steps = 0;
    std::string err;
};

static int bfail(sphgpu_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    return code;
}

#define BCHK(b, call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return bfail(b, SPHGPU_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

template <typename T>
static int balloc(sphgpu_batch* b, T** p, size_t count) {
    void* q = nullptr;
    size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    BCHK(b, hipMalloc(&q, bytes));
    BCHK(b, hipMemsetAsync(q, 0, bytes, b->stream));
    b->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return 0;
}

static void batch_free(sphgpu_batch* b) {
    if (b->stream) hipStreamSynchronize(b->stream);
    for (void* q : b->allocs) hipFree(q);
    b->allocs.clear();
    if (b->h_runs) hipHostFree(b->h_runs);
    if (b->h_err) hipHostFree(b->h_err);
    if (b->dorder) hipFree(b->dorder);
    b->dstage.release();
    for (auto& e : b->ev)
        if (e) hipEventDestroy(e);
    if (b->stream) hipStreamDestroy(b->stream);
}

extern "C" int sphgpu_batch_bytes(int32_t dim, int32_t precision, int64_t slot_capacity,
                                  int64_t cell_capacity, int32_t max_runs, int32_t verlet_cap,
                                  int64_t* bytes) {
    if (!bytes || (dim != 2 && dim != 3) || slot_capacity <= 0 || cell_capacity <= 0 || max_runs <= 0)
        return SPHGPU_E_ARG;
    const int vcap = verlet_cap > 0 ? verlet_cap : (dim == 2 ? kBatchVerletCap2D : kBatchVerletCap3D);
    const int64_t slots = (slot_capacity + kBlock - 1) / kBlock * kBlock;
    *bytes = batch_bytes(dim, precision, slots, cell_capacity, max_runs, vcap);
    return 0;
}

extern "C" int sphgpu_batch_create(int32_t dim, int32_t precision, int64_t slot_capacity,
                                   int64_t cell_capacity, int32_t max_runs, int32_t verlet_cap,
                                   sphgpu_batch** out) {
    if (!out) return create_fail(nullptr, SPHGPU_E_ARG, "out is NULL");
    *out = nullptr;
    if ((dim != 2 && dim != 3) || (precision != SPHGPU_FP64 && precision != SPHGPU_MIXED) ||
        slot_capacity <= 0 || cell_capacity <= 0 || max_runs <= 0 ||
        slot_capacity >= (1LL << 31) - kBlock || cell_capacity >= (1LL << 31) - 2)
        return create_fail(nullptr, SPHGPU_E_ARG, "bad batch dimensions");
    sphgpu_batch* b = new (std::nothrow) sphgpu_batch();
    if (!b) return create_fail(nullptr, SPHGPU_E_ARG, "out of host memory");
    b->dim = dim;
    b->precision = precision;
    b->max_runs = max_runs;
    b->verlet_cap = verlet_cap > 0 ? verlet_cap : (dim == 2 ? kBatchVerletCap2D : kBatchVerletCap3D);
    b->slot_cap = (slot_capacity + kBlock - 1) / kBlock * kBlock;
    b->cell_cap = cell_capacity;
    auto bad = [&](const std::string& why) {
        g_create_error = "sphgpu_batch_create: " + why;
        batch_free(b);
        delete b;
        return SPHGPU_E_HIP;
    };
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { b->stream = nullptr; return bad(hip_why("hipStreamCreateWithFlags", e)); }
    for (auto& ev : b->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess) { ev = nullptr; return bad(hip_why("hipEventCreate", e)); }
    const int64_t N = b->slot_cap;
    Dev& D = b->D;
    int rc = 0;
    rc |= balloc(b, &D.posh, N); rc |= balloc(b, &D.vm, N); rc |= balloc(b, &D.rho, N);
    rc |= balloc(b, &D.e, N); rc |= balloc(b, &D.c, N); rc |= balloc(b, &D.p, N);
    rc |= balloc(b, &D.meta, N); rc |= balloc(b, &D.vprev, N);
    rc |= balloc(b, &D.eprev, N); rc |= balloc(b, &D.div, N); rc |= balloc(b, &D.eta, N);
    rc |= balloc(b, &D.key, N); rc |= balloc(b, &D.refcell, N); rc |= balloc(b, &D.perm, N);
    rc |= balloc(b, &D.cpos, N); rc |= balloc(b, &D.vcnt, N);
    rc |= balloc(b, &D.ecnt, N);
    rc |= balloc(b, &D.vnbr, (size_t)vlist_entries(N, b->verlet_cap));
    if (precision == SPHGPU_MIXED) {
        rc |= balloc(b, &D.dispf, N); rc |= balloc(b, &D.xbf, N);
        rc |= balloc(b, &D.ef, N); rc |= balloc(b, &D.eprevf, N); rc |= balloc(b, &D.divf, N);
        rc |= balloc(b, &D.prf, N); rc |= balloc(b, &D.posf, N);
        rc |= balloc(b, &D.dvf, N); rc |= balloc(b, &D.drho, N); rc |= balloc(b, &D.drho_prev, N);
        rc |= balloc(b, &D.def, N);
    } else {
        rc |= balloc(b, &D.th, N); rc |= balloc(b, &D.de, N); rc |= balloc(b, &D.dv, N);
        rc |= balloc(b, &D.disp, N); rc |= balloc(b, &D.xb, N);
    }
    rc |= balloc(b, &D.partial, (size_t)(N / kBlock + 1) * 6);
    rc |= balloc(b, &D.count, b->cell_cap + 2);
    rc |= balloc(b, &D.start, b->cell_cap + 2);
    rc |= balloc(b, &D.bsum, (b->cell_cap + kScanTile) / kScanTile + 1);
    rc |= balloc(b, &b->d_blk, N / kBlock + 1);
    rc |= balloc(b, &b->d_runof, N);
    rc |= balloc(b, &b->d_runs, max_runs);
    rc |= balloc(b, &b->d_rows, max_runs);
    rc |= balloc(b, &b->d_grid, max_runs);
    rc |= balloc(b, &b->d_vflag, 2 * (size_t)max_runs);
    rc |= balloc(b, &b->d_vany, 2);
    rc |= balloc(b, &b->d_err, 4 * (size_t)max_runs);
    rc |= balloc(b, &b->d_maxcnt, 4 * (size_t)max_runs);
    rc |= balloc(b, &b->d_mat, SPHGPU_MAX_MATERIALS * (size_t)max_runs);
    if (rc) return bad(b->err);
    if ((e = hipHostMalloc((void**)&b->h_runs, sizeof(Params) * max_runs, hipHostMallocDefault)) != hipSuccess) {
        b->h_runs = nullptr;
        return bad(hip_why("hipHostMalloc", e));
    }
    if ((e = hipHostMalloc((void**)&b->h_err, sizeof(int) * 4 * max_runs, hipHostMallocDefault)) != hipSuccess) {
        b->h_err = nullptr;
        return bad(hip_why("hipHostMalloc", e));
    }
    std::memset(b->h_runs, 0, sizeof(Params) * max_runs);
    // every slot starts free
    std::vector<int> neg(N, -1);
    if ((e = hipMemcpy(b->d_runof, neg.data(), N * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
        return bad(hip_why("hipMemcpy", e));
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return bad(hip_why("hipStreamSynchronize", e));
    b->slots.init(N);
    b->cells.init(b->cell_cap);
    b->runs.resize(max_runs);
    *out = b;
    return 0;
}

// Mark a run's slots as its own (row) or free (-1).
static int set_runof(sphgpu_batch* b, const BatchRun& r, int row) {
    std::vector<int> v(r.nslot, -1);
    for (int64_t k = 0; k < r.n; ++k) v[k] = row;
    BCHK(b, hipMemcpyAsync(b->d_runof + r.s0, v.data(), r.nslot * sizeof(int), hipMemcpyHostToDevice,
                           b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    return 0;
}

extern "C" int sphgpu_batch_add(sphgpu_batch* b, const sphgpu_config* cfg, int64_t n_real,
                                int64_t n_dummy, const double* x, const double* v, const double* mass,
                                const double* rho, const double* p, const double* e, const double* c,
                                const double* h, const int32_t* type, const int32_t* state,
                                const int32_t* boundary, int64_t nsteps, int32_t win,
                                double skin_frac, int32_t* run) {
    if (!b || !cfg || !run) return bfail(b, SPHGPU_E_ARG, "NULL argument");
    if (nsteps < 0) return bfail(b, SPHGPU_E_ARG, "negative step count");
    std::string why;
    int rc = validate_config(cfg, &why);
    if (rc) return bfail(b, rc, why);
    if (cfg->dim != b->dim || cfg->precision != b->precision)
        return bfail(b, SPHGPU_E_BAD_CONFIG, "run dimension or precision differs from the batch");
    Tuning tune = env_tuning();
    if (win >= 1 && win <= 4) tune.win = win;
    if (skin_frac > 0) tune.skin_frac = skin_frac;
    HostRun hr;
    rc = prepare_run(*cfg, n_real, n_dummy, x, v, mass, rho, p, e, c, h, nullptr, type, state, boundary,
                     true, false, tune, hr, why);
    if (rc) return bfail(b, rc, why);
    if (!hr.fast)
        return bfail(b, SPHGPU_E_BAD_CONFIG,
                     "batched runs use the fluid fast path (cubic spline, continuity density, "
                     "no physical viscosity, solids, kernel correction or artificial stress)");
    if (hr.P.pair_ext)
        return bfail(b, SPHGPU_E_BAD_CONFIG, "batched runs have no wall-force range or second fluid");
    int row = -1;
    for (int k = 0; k < b->max_runs; ++k)
        if (!b->runs[k].used) { row = k; break; }
    if (row < 0) return bfail(b, SPHGPU_E_CAPACITY, "run table full");
    const int64_t nslot = (hr.n + kBlock - 1) / kBlock * kBlock;
    const int64_t ncells = hr.P.ncell + 1;
    const int64_t s0 = b->slots.alloc(nslot);
    if (s0 < 0) return bfail(b, SPHGPU_E_CAPACITY, "no free slot range for this run");
    const int64_t c0 = b->cells.alloc(ncells);
    if (c0 < 0) {
        b->slots.release(s0);
        return bfail(b, SPHGPU_E_CAPACITY, "no free cell range for this run");
    }
    BatchRun& r = b->runs[row];
    r = BatchRun{};
    r.used = true;
    r.runnable = nsteps > 0;
    r.cfg = *cfg;
    r.n = hr.n;
    r.n_real = hr.n_real;
    r.type_idx = std::move(hr.type_idx);
    r.s0 = s0;
    r.nslot = nslot;
    r.c0 = c0;
    r.ncells = ncells;
    r.step = 1;
    r.target = nsteps;
    Params& P = r.P;
    P = hr.P;
    P.s0 = (int)s0;
    P.c0 = c0;
    P.nblk = (int)((hr.n + kBlock - 1) / kBlock);
    P.step0 = 1;
    P.grid = b->d_grid + row;
    P.vflag = b->d_vflag + 2 * row;
    P.vany = b->d_vany;
    P.err = b->d_err + 4 * row;
    P.maxcnt = b->d_maxcnt + 4 * row;
    P.partial = b->D.partial + 6 * (s0 / kBlock);
    P.mat = b->d_mat + SPHGPU_MAX_MATERIALS * row;
    P.verlet_cap = b->verlet_cap;
    P.nbr_cap = 0;
    Material tab[SPHGPU_MAX_MATERIALS];
    for (int k = 0; k < SPHGPU_MAX_MATERIALS; ++k) {
        const sphgpu_material& m = cfg->mat[k];
        tab[k] = Material{m.type, m.eos, m.rho0, m.K, m.gamma, m.eta, m.sigma_y};
    }
    const int64_t n = hr.n;
    Dev& D = b->D;
    auto undo = [&](int code, const std::string& m) {
        b->slots.release(s0);
        b->cells.release(c0);
        r = BatchRun{};
        return bfail(b, code, m);
    };
#define ACHK(call)                                                                      \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) return undo(SPHGPU_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)
    ACHK(hipMemcpyAsync(b->d_mat + SPHGPU_MAX_MATERIALS * row, tab, sizeof(tab), hipMemcpyHostToDevice, b->stream));
    ACHK(hipMemcpyAsync(D.posh + s0, hr.posh.data(), n * sizeof(double4), hipMemcpyHostToDevice, b->stream));
    ACHK(hipMemcpyAsync(D.vm + s0, hr.vm.data(), n * sizeof(double4), hipMemcpyHostToDevice, b->stream));
    ACHK(hipMemcpyAsync(D.rho + s0, hr.rho.data(), n * sizeof(double), hipMemcpyHostToDevice, b->stream));
    ACHK(hipMemcpyAsync(D.e + s0, hr.e.data(), n * sizeof(double), hipMemcpyHostToDevice, b->stream));
    ACHK(hipMemcpyAsync(D.c + s0, hr.c.data(), n * sizeof(double), hipMemcpyHostToDevice, b->stream));
    ACHK(hipMemcpyAsync(D.p + s0, hr.p.data(), n * sizeof(double), hipMemcpyHostToDevice, b->stream));
    ACHK(hipMemcpyAsync(D.meta + s0, hr.meta.data(), n * sizeof(int4), hipMemcpyHostToDevice, b->stream));
    ACHK(hipMemsetAsync(D.div + s0, 0, nslot * sizeof(double2), b->stream));
    ACHK(hipMemsetAsync(D.ecnt + s0, 0, nslot * sizeof(int), b->stream));
    if (b->precision == SPHGPU_MIXED) {
        ACHK(hipMemcpyAsync(D.drho + s0, hr.drho.data(), n * sizeof(float), hipMemcpyHostToDevice, b->stream));
        ACHK(hipMemsetAsync(D.dvf + s0, 0, nslot * sizeof(float4), b->stream));
        ACHK(hipMemsetAsync(D.dispf + s0, 0, nslot * sizeof(float4), b->stream));
        ACHK(hipMemsetAsync(D.divf + s0, 0, nslot * sizeof(float2), b->stream));
        {
            std::vector<float> ef(hr.e.begin(), hr.e.end());
            ACHK(hipMemcpyAsync(D.ef + s0, ef.data(), n * sizeof(float), hipMemcpyHostToDevice, b->stream));
            ACHK(hipStreamSynchronize(b->stream));
        }
        ACHK(hipMemsetAsync(D.def + s0, 0, nslot * sizeof(float), b->stream));
    } else {
        ACHK(hipMemsetAsync(D.dv + s0, 0, nslot * sizeof(double4), b->stream));
        ACHK(hipMemsetAsync(D.disp + s0, 0, nslot * sizeof(double4), b->stream));
        ACHK(hipMemsetAsync(D.de + s0, 0, nslot * sizeof(double), b->stream));
    }
    // no candidate list yet: force the run's first build, clear its error words
    const unsigned long long force[2] = {1ull, 0ull};
    ACHK(hipMemcpyAsync(P.vflag, force, sizeof(force), hipMemcpyHostToDevice, b->stream));
    ACHK(hipMemcpyAsync(b->d_vany, force, sizeof(unsigned long long), hipMemcpyHostToDevice, b->stream));
    ACHK(hipMemsetAsync(P.err, 0, 4 * sizeof(int), b->stream));
    ACHK(hipMemsetAsync(P.maxcnt, 0, 4 * sizeof(int), b->stream));
#undef ACHK
    rc = set_runof(b, r, r.runnable ? row : -1);
    if (rc) return undo(rc, b->err);
    b->map_dirty = true;
    *run = row;
    return 0;
}

// Launch map of the runnable runs (ascending slot order keeps neighbouring
// runs on neighbouring blocks).
static int rebuild_map(sphgpu_batch* b) {
    std::vector<int> order;
    for (int k = 0; k < b->max_runs; ++k)
        if (b->runs[k].used && b->runs[k].runnable) order.push_back(k);
    std::sort(order.begin(), order.end(), [&](int a, int c) { return b->runs[a].s0 < b->runs[c].s0; });
    b->blk.clear();
    b->rows = order;
    b->nact = 0;
    for (int k : order) {
        const BatchRun& r = b->runs[k];
        for (int q = 0; q < r.P.nblk; ++q) b->blk.push_back(make_int2(k, q));
        b->nact += (int)r.n;
    }
    if (!b->blk.empty())
        BCHK(b, hipMemcpyAsync(b->d_blk, b->blk.data(), b->blk.size() * sizeof(int2), hipMemcpyHostToDevice,
                               b->stream));
    if (!b->rows.empty())
        BCHK(b, hipMemcpyAsync(b->d_rows, b->rows.data(), b->rows.size() * sizeof(int), hipMemcpyHostToDevice,
                               b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    b->map_dirty = false;
    return 0;
}

template <int DIM, bool MIXED>
static void batch_chunk(sphgpu_batch* b, int m) {
    const int mc = (int)b->cells.high_water();
    const RunMap M{b->d_runs, b->d_blk, b->d_runof, b->d_rows, (int)b->slots.high_water(), b->nact};
    const Params P0{};
    Dev& D = b->D;
    hipStream_t st = b->stream;
    const int nbt = (int)b->blk.size();
    const int nr = (int)b->rows.size();
    const int nss = std::max(1, std::min((M.nslot + kBlock - 1) / kBlock, kStrideBlocks));
    const int sb = (mc + kScanTile - 1) / kScanTile;
    const Gate G{b->d_vany};
    hipLaunchKernelGGL(k_bbox<true>, dim3(nbt), dim3(kBlock), 0, st, P0, D, M);
    hipLaunchKernelGGL((k_grid<DIM, true>), dim3(nr), dim3(kGridThreads), 0, st, P0, D, M);
    hipLaunchKernelGGL((k_point_fast<DIM, MIXED, true>), dim3(nbt), dim3(kBlock), 0, st, P0, D, 0, M);
    for (int j = 0; j < m; ++j) {
        hipLaunchKernelGGL((k_keys<DIM, true>), dim3(nss), dim3(kBlock), 0, st, P0, D, G, 0, M);
        const int own = sb <= kOwnPrefixTiles;
        hipLaunchKernelGGL(k_scan1, dim3(sb), dim3(kBlock), 0, st, D.count, mc, D.bsum, G);
        if (!own) hipLaunchKernelGGL(k_scan2, dim3(1), dim3(kBlock), 0, st, D.bsum, sb, G);
        hipLaunchKernelGGL(k_scan3, dim3(sb), dim3(kBlock), 0, st, D.count, mc, D.bsum, D.start, own, G);
        hipLaunchKernelGGL(k_scatter<true>, dim3(nss), dim3(kBlock), 0, st, P0, D, G, M);
        hipLaunchKernelGGL(k_cellsort<true>, dim3(std::max(1, std::min((mc + kBlock - 1) / kBlock, kStrideBlocks))),
                           dim3(kBlock), 0, st, P0, D, mc, G, M);
        hipLaunchKernelGGL((k_verlet_build<DIM, true>), dim3(nbt), dim3(kBlock), 0, st, P0, D, Gate{nullptr}, 0, M);
        hipLaunchKernelGGL((k_force_fused<DIM, MIXED, false, true>), dim3(nbt), dim3(kBlock), 0, st, P0, D, M);
        if (j + 1 < m) {
            hipLaunchKernelGGL((k_correct_predict<DIM, MIXED, true>), dim3(nbt), dim3(kBlock), 0, st, P0, D, 0, M, j);
            hipLaunchKernelGGL((k_grid<DIM, true>), dim3(nr), dim3(kGridThreads), 0, st, P0, D, M);
            hipLaunchKernelGGL((k_refcell<DIM, true>), dim3(nss), dim3(kBlock), 0, st, P0, D, M);
        } else {
            hipLaunchKernelGGL((k_correct<DIM, MIXED, true>), dim3(nbt), dim3(kBlock), 0, st, P0, D, 0, M, j);
        }
    }
}

static int batch_errors(sphgpu_batch* b) {
    BCHK(b, hipMemcpyAsync(b->h_err, b->d_err, sizeof(int) * 4 * b->max_runs, hipMemcpyDeviceToHost, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    for (int k = 0; k < b->max_runs; ++k) {
        BatchRun& r = b->runs[k];
        if (!r.used || !r.runnable) continue;
        const int e = b->h_err[4 * k];
        if (!e) continue;
        r.runnable = false;
        if (e & kErrNonFinite) { r.status = SPHGPU_E_NONFINITE; r.msg = "non-finite position, velocity or density"; }
        else if (e & kErrOutOfRange) { r.status = SPHGPU_E_OUT_OF_RANGE; r.msg = "particle out of range"; }
        else { r.status = SPHGPU_E_PAIR_CAPACITY; r.msg = "candidate list capacity exceeded"; }
        int rc = set_runof(b, r, -1);
        if (rc) return rc;
        b->map_dirty = true;
    }
    return 0;
}

extern "C" int sphgpu_batch_advance(sphgpu_batch* b, int32_t nsteps) {
    if (!b) return SPHGPU_E_ARG;
    if (nsteps < 0) return bfail(b, SPHGPU_E_ARG, "negative step count");
    b->last_steps = 0;
    b->last_particle_steps = 0;
    BCHK(b, hipEventRecord(b->ev[0], b->stream));
    int left = nsteps;
    while (left > 0) {
        int64_t m = left;
        bool any = false;
        for (auto& r : b->runs)
            if (r.used && r.runnable) {
                any = true;
                m = std::min<int64_t>(m, r.target - r.step + 1);
            }
        if (!any) break;
        if (b->map_dirty) {
            int rc = rebuild_map(b);
            if (rc) return rc;
        }
        for (int k = 0; k < b->max_runs; ++k) {
            BatchRun& r = b->runs[k];
            if (!r.used) continue;
            r.P.step0 = r.step;
            b->h_runs[k] = r.P;
        }
        BCHK(b, hipMemcpyAsync(b->d_runs, b->h_runs, sizeof(Params) * b->max_runs, hipMemcpyHostToDevice,
                               b->stream));
        if (b->dim == 2) {
            if (b->precision == SPHGPU_MIXED) batch_chunk<2, true>(b, (int)m);
            else batch_chunk<2, false>(b, (int)m);
        } else {
            if (b->precision == SPHGPU_MIXED) batch_chunk<3, true>(b, (int)m);
            else batch_chunk<3, false>(b, (int)m);
        }
        BCHK(b, hipGetLastError());
        for (auto& r : b->runs)
            if (r.used && r.runnable) {
                r.step += m;
                b->last_particle_steps += r.n_real * m;
            }
        b->last_steps += m;
        int rc = batch_errors(b);   // also waits for the chunk (the staging buffer is reused)
        if (rc) return rc;
        for (int k = 0; k < b->max_runs; ++k) {
            BatchRun& r = b->runs[k];
            if (r.used && r.runnable && r.step > r.target) {
                r.runnable = false;
                rc = set_runof(b, r, -1);
                if (rc) return rc;
                b->map_dirty = true;
            }
        }
        left -= (int)m;
    }
    BCHK(b, hipEventRecord(b->ev[1], b->stream));
    BCHK(b, hipEventSynchronize(b->ev[1]));
    float ms = 0;
    hipEventElapsedTime(&ms, b->ev[0], b->ev[1]);
    b->last_ms = ms;
    return 0;
}

static BatchRun* batch_run(sphgpu_batch* b, int32_t run) {
    if (!b || run < 0 || run >= b->max_runs || !b->runs[run].used) return nullptr;
    return &b->runs[run];
}

extern "C" int sphgpu_batch_status(sphgpu_batch* b, int32_t run, int64_t* next_step, int32_t* state) {
    BatchRun* r = batch_run(b, run);
    if (!r) return bfail(b, SPHGPU_E_ARG, "no such run");
    if (next_step) *next_step = r->step;
    if (state) *state = r->status ? r->status : (r->runnable ? 0 : 1);
    return 0;
}

extern "C" const char* sphgpu_batch_run_error(sphgpu_batch* b, int32_t run) {
    BatchRun* r = batch_run(b, run);
    return r ? r->msg.c_str() : "no such run";
}

extern "C" int sphgpu_batch_download(sphgpu_batch* b, int32_t run, int32_t field, double* dst) {
    BatchRun* r = batch_run(b, run);
    if (!r) return bfail(b, SPHGPU_E_ARG, "no such run");
    if (!dst || field < 0 || field >= SPHGPU_F_NFIELDS) return bfail(b, SPHGPU_E_ARG, "bad field");
    // a view of the arena that starts at the run's first slot
    Dev V = b->D;
    const int64_t s0 = r->s0;
    auto off = [&](auto*& p) { if (p) p += s0; };
    off(V.posh); off(V.vm); off(V.rho); off(V.e); off(V.c); off(V.p); off(V.meta); off(V.disp);
    off(V.dv); off(V.de); off(V.div); off(V.eta); off(V.drho); off(V.dvf); off(V.dispf);
    off(V.ef); off(V.divf);
    off(V.def);
    // device scratch for the reorder: the widest field of this run (downloads
    // synchronise, so nothing in flight uses the old buffer)
    const size_t need = (size_t)r->n * b->dim * b->dim * sizeof(double);
    if (need > b->dorder_bytes) {
        if (b->dorder) hipFree(b->dorder);
        b->dorder = nullptr;
        b->dorder_bytes = 0;
        hipError_t e = hipMalloc(&b->dorder, need);
        if (e != hipSuccess) {
            b->dorder = nullptr;
            return bfail(b, SPHGPU_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        b->dorder_bytes = need;
    }
    const bool typed = (int64_t)r->type_idx.size() == r->n;
    std::string msg;
    int rc = download_view(V, r->n, b->dim, b->precision == SPHGPU_MIXED, r->cfg, b->stream, field, dst, msg,
                           typed ? b->dorder : nullptr, typed ? r->type_idx.data() : nullptr, &b->dstage,
                           r->P.err + 3);
    return rc ? bfail(b, rc, msg) : 0;
}

extern "C" int sphgpu_batch_remove(sphgpu_batch* b, int32_t run) {
    BatchRun* r = batch_run(b, run);
    if (!r) return bfail(b, SPHGPU_E_ARG, "no such run");
    BCHK(b, hipStreamSynchronize(b->stream));
    if (r->runnable) {
        int rc = set_runof(b, *r, -1);
        if (rc) return rc;
    }
    b->slots.release(r->s0);
    b->cells.release(r->c0);
    *r = BatchRun{};
    b->map_dirty = true;
    return 0;
}

extern "C" int sphgpu_batch_info(sphgpu_batch* b, int64_t* out8) {
    if (!b || !out8) return SPHGPU_E_ARG;
    int64_t used = 0, runnable = 0;
    for (auto& r : b->runs) {
        used += r.used;
        runnable += r.used && r.runnable;
    }
    out8[0] = used;
    out8[1] = runnable;
    out8[2] = b->slots.in_use();
    out8[3] = b->slot_cap;
    out8[4] = b->cells.in_use();
    out8[5] = b->cell_cap;
    out8[6] = b->max_runs;
    out8[7] = b->verlet_cap;
    return 0;
}

extern "C" int sphgpu_batch_timing(sphgpu_batch* b, double* ms, int64_t* steps, int64_t* particle_steps) {
    if (!b) return SPHGPU_E_ARG;
    if (ms) *ms = b->last_ms;
    if (steps) *steps = b->last_steps;
    if (particle_steps) *particle_steps = b->last_particle_steps;
    return 0;
}

extern "C" const char* sphgpu_batch_last_error(sphgpu_batch* b) {
    if (b) return b->err.c_str();
    return g_create_error.empty() ? "no batch" : g_create_error.c_str();
}

extern "C" void sphgpu_batch_destroy(sphgpu_batch* b) {
    if (!b) return;
    batch_free(b);
    delete b;
}
