// Exact checkpoints of a single-run context (part of sphgpu.hip).
//
// A checkpoint holds every device allocation of the context byte for byte
// (state, integrator history, slot order, cell tables, candidate list,
// rebuild gate) and the host-side step state. A context restored from it
// takes the same kernel sequence on the same data, so it continues bit for
// bit where the saved one would have. The reference's nearest counterpart
// is its particle file (read_initial_file / write_initial_file,
// input.f90:22-92, through the derived-type I/O of SPH.f90:56-93), which
// holds the particle fields but not the integrator history.
//
// Blob: header | config | Params | Dev | per allocation (old base, bytes) |
// the allocations' bytes, each padded to 8. Device pointers inside Params and
// Dev are stored as they were and moved onto the new allocations at restore.

namespace {

constexpr uint64_t kCkptMagic = 0x31544b4350485053ull;   // "SPHPCKT1", little-endian
constexpr uint32_t kCkptVersion = 1;

struct CkptHeader {
    uint64_t magic;
    uint32_t version, abi;
    uint32_t cfg_bytes, params_bytes, dev_bytes, nalloc;
    int64_t capacity, n, n_real, n_dummy;
    int64_t ncell_alloc, nbr_alloc, vnbr_alloc, step;
    int32_t dim, nblk;
    int32_t has_state, use_verlet, fused_done, force_general;
};

inline size_t pad8(size_t b) { return (b + 7) & ~(size_t)7; }

static_assert(sizeof(Dev) % sizeof(void*) == 0, "Dev holds device pointers only");

size_t ckpt_bytes(const sphgpu_ctx* c) {
    size_t b = sizeof(CkptHeader) + pad8(sizeof(sphgpu_config)) + pad8(sizeof(Params)) +
               pad8(sizeof(Dev)) + c->allocs.size() * 2 * sizeof(uint64_t);
    for (size_t k = 0; k < c->allocs.size(); ++k) b += pad8(c->alloc_bytes[k]);
    return b;
}

// Bytes from a saved device address to the end of the saved allocation that
// holds it (0 when none does).
inline uint64_t bytes_after(const void* p, const std::vector<uint64_t>& old_base,
                            const std::vector<uint64_t>& bytes) {
    const uint64_t a = (uint64_t)(uintptr_t)p;
    for (size_t k = 0; k < old_base.size(); ++k)
        if (a >= old_base[k] && a < old_base[k] + bytes[k]) return old_base[k] + bytes[k] - a;
    return 0;
}

// Map a device address of the saved context onto the restored allocations.
template <typename T>
bool remap(T*& p, const std::vector<uint64_t>& old_base, const std::vector<uint64_t>& bytes,
           const std::vector<void*>& fresh) {
    if (!p) return true;
    const uint64_t a = (uint64_t)(uintptr_t)p;
    for (size_t k = 0; k < old_base.size(); ++k)
        if (a >= old_base[k] && a < old_base[k] + bytes[k]) {
            p = (T*)((char*)fresh[k] + (a - old_base[k]));
            return true;
        }
    return false;
}

}  // namespace

extern "C" int sphgpu_context_info(sphgpu_ctx* c, sphgpu_config* cfg_out, int64_t* counts3) {
    if (!c || !counts3) return fail(c, SPHGPU_E_ARG, "NULL context or output");
    if (cfg_out) *cfg_out = c->cfg;
    counts3[0] = c->n_real;
    counts3[1] = c->n_dummy;
    counts3[2] = c->capacity;
    return 0;
}

extern "C" int sphgpu_checkpoint_size(sphgpu_ctx* c, int64_t* bytes) {
    if (!c || !bytes) return fail(c, SPHGPU_E_ARG, "NULL context or output");
    if (!c->has_state) return fail(c, SPHGPU_E_ARG, "no particle state uploaded");
    *bytes = (int64_t)ckpt_bytes(c);
    return 0;
}

extern "C" int sphgpu_checkpoint_save(sphgpu_ctx* c, void* buf, int64_t bytes) {
    if (!c || !buf) return fail(c, SPHGPU_E_ARG, "NULL context or buffer");
    if (!c->has_state) return fail(c, SPHGPU_E_ARG, "no particle state uploaded");
    if (bytes < (int64_t)ckpt_bytes(c))
        return fail(c, SPHGPU_E_CAPACITY, "checkpoint buffer smaller than sphgpu_checkpoint_size");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    char* o = (char*)buf;
    CkptHeader h{};
    h.magic = kCkptMagic;
    h.version = kCkptVersion;
    h.abi = SPHGPU_ABI_VERSION;
    h.cfg_bytes = sizeof(sphgpu_config);
    h.params_bytes = sizeof(Params);
    h.dev_bytes = sizeof(Dev);
    h.nalloc = (uint32_t)c->allocs.size();
    h.capacity = c->capacity;
    h.n = c->n;
    h.n_real = c->n_real;
    h.n_dummy = c->n_dummy;
    h.ncell_alloc = c->ncell_alloc;
    h.nbr_alloc = c->nbr_alloc;
    h.vnbr_alloc = c->vnbr_alloc;
    h.step = c->step;
    h.dim = c->dim;
    h.nblk = c->nblk;
    h.has_state = c->has_state;
    h.use_verlet = c->use_verlet;
    h.fused_done = c->fused_done;
    h.force_general = c->force_general;
    memcpy(o, &h, sizeof(h));
    o += sizeof(h);
    memcpy(o, &c->cfg, sizeof(sphgpu_config));
    o += pad8(sizeof(sphgpu_config));
    memcpy(o, &c->P, sizeof(Params));
    o += pad8(sizeof(Params));
    memcpy(o, &c->D, sizeof(Dev));
    o += pad8(sizeof(Dev));
    for (size_t k = 0; k < c->allocs.size(); ++k) {
        const uint64_t rec[2] = {(uint64_t)(uintptr_t)c->allocs[k], (uint64_t)c->alloc_bytes[k]};
        memcpy(o, rec, sizeof(rec));
        o += sizeof(rec);
    }
    for (size_t k = 0; k < c->allocs.size(); ++k) {
        HIPCHK(c, hipMemcpy(o, c->allocs[k], c->alloc_bytes[k], hipMemcpyDeviceToHost));
        o += pad8(c->alloc_bytes[k]);
    }
    return 0;
}

extern "C" int sphgpu_checkpoint_restore(const void* buf, int64_t bytes, sphgpu_ctx** out) {
    if (!out) return create_fail(nullptr, SPHGPU_E_ARG, "out is NULL");
    *out = nullptr;
    if (!buf || bytes < (int64_t)sizeof(CkptHeader))
        return create_fail(nullptr, SPHGPU_E_ARG, "NULL or short checkpoint");
    const char* in = (const char*)buf;
    CkptHeader h;
    memcpy(&h, in, sizeof(h));
    if (h.magic != kCkptMagic || h.version != kCkptVersion)
        return create_fail(nullptr, SPHGPU_E_ARG, "not a checkpoint of this library");
    if (h.abi != SPHGPU_ABI_VERSION || h.cfg_bytes != sizeof(sphgpu_config) ||
        h.params_bytes != sizeof(Params) || h.dev_bytes != sizeof(Dev))
        return create_fail(nullptr, SPHGPU_E_ARG, "checkpoint written by another library version");
    const char* p = in + sizeof(h);
    sphgpu_config cfg;
    memcpy(&cfg, p, sizeof(cfg));
    p += pad8(sizeof(sphgpu_config));
    Params P;
    memcpy(&P, p, sizeof(P));
    p += pad8(sizeof(Params));
    Dev D;
    memcpy(&D, p, sizeof(D));
    p += pad8(sizeof(Dev));
    std::vector<uint64_t> old_base(h.nalloc), nbytes(h.nalloc);
    size_t need = (size_t)(p - in) + (size_t)h.nalloc * 2 * sizeof(uint64_t);
    if ((int64_t)need > bytes) return create_fail(nullptr, SPHGPU_E_ARG, "truncated checkpoint");
    for (uint32_t k = 0; k < h.nalloc; ++k) {
        uint64_t rec[2];
        memcpy(rec, p, sizeof(rec));
        p += sizeof(rec);
        old_base[k] = rec[0];
        nbytes[k] = rec[1];
        need += pad8(rec[1]);
    }
    if ((int64_t)need > bytes) return create_fail(nullptr, SPHGPU_E_ARG, "truncated checkpoint");
    // what the header and the saved config say must agree with each other and
    // with the allocations before anything is created on the device: a
    // corrupted blob would otherwise send kernels past the ends of the arrays
    {
        std::string why;
        if (validate_config(&cfg, &why)) return create_fail(nullptr, SPHGPU_E_ARG, "checkpoint config: " + why);
        const bool counts_ok = h.capacity > 0 && h.n >= 1 && h.n <= h.capacity && h.n_real >= 0 &&
                               h.n_dummy >= 0 && h.n_real + h.n_dummy == h.n && h.n <= INT32_MAX;
        if (!counts_ok) return create_fail(nullptr, SPHGPU_E_ARG, "checkpoint header: inconsistent particle counts");
        if (h.dim != cfg.dim || (h.dim != 2 && h.dim != 3))
            return create_fail(nullptr, SPHGPU_E_ARG, "checkpoint header: dim does not match the config");
        if (P.n != h.n || P.n_real != h.n_real || P.dim != h.dim)
            return create_fail(nullptr, SPHGPU_E_ARG, "checkpoint header: counts do not match the run's parameters");
        if (P.nblk != (int)((h.n + kBlock - 1) / kBlock) || h.step < 1)
            return create_fail(nullptr, SPHGPU_E_ARG, "checkpoint: block count or step out of range");
        // every per-particle array must hold n elements from where it starts
        const struct { const void* p; uint64_t elem; } arrays[] = {
            {D.posh, sizeof(double4)}, {D.vm, sizeof(double4)}, {D.rho, sizeof(double)},
            {D.e, sizeof(double)}, {D.c, sizeof(double)}, {D.p, sizeof(double)},
            {D.meta, sizeof(int4)}, {D.dv, sizeof(double4)}, {D.de, sizeof(double)},
            {D.vprev, sizeof(double4)}, {D.th, sizeof(double4)}, {D.key, sizeof(uint32_t)},
            {D.perm, sizeof(int)}, {D.cpos, sizeof(float4)}};
        for (const auto& a : arrays)
            if (!a.p || bytes_after(a.p, old_base, nbytes) < (uint64_t)h.n * a.elem)
                return create_fail(nullptr, SPHGPU_E_ARG, "checkpoint: a particle array is smaller than its count");
    }
    sphgpu_ctx* c = nullptr;
    int rc = sphgpu_create(&cfg, h.capacity, &c);
    if (rc) return rc;
    c->n = h.n;
    c->n_real = h.n_real;
    c->n_dummy = h.n_dummy;
    c->dim = h.dim;
    c->nblk = h.nblk;
    c->step = h.step;
    c->use_verlet = h.use_verlet;
    c->fused_done = h.fused_done;
    c->force_general = h.force_general;
    c->ncell_alloc = h.ncell_alloc;
    c->nbr_alloc = h.nbr_alloc;
    c->vnbr_alloc = h.vnbr_alloc;
    auto bail = [&](int code, const std::string& why) {
        sphgpu_destroy(c);
        return create_fail(nullptr, code, why);
    };
    for (uint32_t k = 0; k < h.nalloc; ++k) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, nbytes[k]);
        if (e != hipSuccess) return bail(SPHGPU_E_HIP, hip_why("hipMalloc", e));
        c->allocs.push_back(q);
        c->alloc_bytes.push_back(nbytes[k]);
        e = hipMemcpy(q, p, nbytes[k], hipMemcpyHostToDevice);
        if (e != hipSuccess) return bail(SPHGPU_E_HIP, hip_why("hipMemcpy", e));
        p += pad8(nbytes[k]);
    }
    // Dev holds device pointers only
    void** dp = (void**)&D;
    bool ok = true;
    for (size_t k = 0; k < sizeof(Dev) / sizeof(void*); ++k)
        ok &= remap(dp[k], old_base, nbytes, c->allocs);
    ok &= remap(P.mat, old_base, nbytes, c->allocs);
    ok &= remap(P.grid, old_base, nbytes, c->allocs);
    ok &= remap(P.vflag, old_base, nbytes, c->allocs);
    ok &= remap(P.vany, old_base, nbytes, c->allocs);
    ok &= remap(P.err, old_base, nbytes, c->allocs);
    ok &= remap(P.maxcnt, old_base, nbytes, c->allocs);
    ok &= remap(P.partial, old_base, nbytes, c->allocs);
    if (!ok) return bail(SPHGPU_E_ARG, "checkpoint holds a device address outside its allocations");
    c->D = D;
    c->P = P;
    c->has_state = h.has_state != 0;
    *out = c;
    return 0;
}
