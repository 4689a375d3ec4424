// Candidate list for the general path of libsphgpu (included by sphgpu.hip
// after the fluid fast path, whose list build it shares): the solid, KGC,
// summation-density, SLE and SPH 1/3 physics keep their pair kernels over an
// exact neighbour list, but that list is no longer found by a cell sweep every
// step. The cell sort and the candidate list (every j within cut + skin) are
// rebuilt only when some particle has moved skin/2 since the last build, and
// each step filters the candidates with the reference's exact rule into the
// exact list.
//
// Exactness is the fast path's argument: a pair within the cutoff at this
// step was within cutoff + skin at the build, so it is a candidate, and the
// filter accepts exactly the reference's pairs (asymmetricSearch.f90:209-237).
// With an evolving h the list is built for the planned largest h, which the
// smoothing-length update enforces (k_hsml).

// Predict + EOS (as k_point_build), the FP32 position record of the filter's
// pre-test (and in mixed precision the FP32 pair record), the reference
// cells of this step when the reference window can exclude a pair, and the
// displacement vote for a rebuild. mode bit 0: first step (no predict); bit
// 1: no predict at all; bit 2: sizing pass (the state is left as uploaded).
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_point_gen(Params P, Dev D, int mode) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    bool moved = false;
    if (s < P.n) {
        const int4 mt = D.meta[s];
        const bool real = mt.w < P.n_real;
        if (real && !(mode & 2)) predict_one<DIM>(P, D, s, mt, mode & 1);
        if (!(mode & 4)) {
            if (P.pa_sph == 3) D.pold[s] = D.p[s];
            if (!P.summation_density) eos_th(P, D, s, mt);
        }
        const double4 q = D.posh[s];
        D.posf[s] = cell_offset_record(P, q);
        if (P.pair32) write_rec32(D, s);
        if (mt.y != -1 && P.grid->need_win) {
            const RefGrid g = *P.grid;
            bool bad = false;
            const double xs[3] = {q.x, q.y, q.z};
            int rc[3] = {1, 1, 1};
            for (int d = 0; d < DIM; ++d) rc[d] = ref_locate(xs[d], g.minc[d], g.maxc[d], g.num[d], &bad);
            if (bad) atomicOr(P.err, (isfinite(q.x) && isfinite(q.y) && isfinite(q.z)) ? kErrOutOfRange
                                                                                      : kErrNonFinite);
            D.refcell[s] = pack_ref(rc[0], rc[1], rc[2]);
        }
        moved = moved_half_skin<false>(P, D, s, q);
    }
    // every lane reaches the vote (one atomic per wave)
    vote_rebuild(P, moved);
}

// Exact neighbour list of target s for this step from its candidate list:
// the reference's distance test and window on every candidate, kept in
// candidate order (the order of the cell sweep at the build), in the
// wave-blocked layout the pair kernels read. The distance is tested in FP32
// from the cell-offset records first; only a candidate within the FP32 error
// bound of the cutoff (or off the grid) loads its FP64 position for the
// reference's test. A particle whose candidates outgrew its list walks its
// cell window of the last sort instead. visit(j, pf_j) is called for every
// accepted neighbour in list order (pf_j its cell-offset record). Returns the
// neighbour count (the list holds at most nbr_cap of them).
//
// LPT > 1 (small runs, no visitor): the LPT lanes of a group (sub = 0..LPT-1)
// test LPT consecutive candidates at once; a ballot over the group gives each
// accepted one its place, so the list is the same as one lane's. A particle
// whose candidates overflowed walks its cell window on lane 0 of its group.
// The returned count is lane 0's.
template <int DIM, int LPT = 1, typename V>
__device__ __forceinline__ int filter_target(const Params& P, const Dev& D, int s, const double4& qi,
                                             V&& visit, int sub = 0) {
    const double cut = qi.w * (double)P.scale_k;
    const double cut2 = cut * cut;
    const double lo2 = cut2 * (1.0 - 1e-12), hi2 = cut2 * (1.0 + 1e-12);
    // FP32 r^2 from cell offsets is within M of the FP64 one
    const float4 pfi = D.posf[s];
    const uint32_t pk_i = __float_as_uint(pfi.w);
    int cxi, cyi, czi;
    unpack_cell<DIM>(pk_i, cxi, cyi, czi);
    const float csf = (float)P.cs;
    float lof, hif;
    {
        const double u = 5.9604644775390625e-8;
        const double e = 2.0 * u * P.cs + u * cut;
        const double M = 2.0 * 1.7320508075688772 * cut * e + 3.0 * e * e + 3.0 * u * cut2;
        lof = (float)(cut2 - 2.0 * M - 1e-6 * cut2);
        hif = (float)(cut2 + 2.0 * M + 1e-6 * cut2);
    }
    // reference window (only when it can exclude a pair of this particle)
    const RefGrid g = *P.grid;
    int wr[3] = {0, 0, 0}, rc[3] = {0, 0, 0};
    bool need_win = false;
    const uint32_t rci = g.need_win ? D.refcell[s] : 0u;
    for (int d = 0; d < DIM && g.need_win; ++d) {
        const double span = g.maxc[d] - g.minc[d];
        wr[d] = (int)(qi.w / span * g.num[d]) + 1;   // cellNumPerHSML, asymmetricSearch.f90:209-211
        rc[d] = ref_axis(rci, d);
        const double reach = cut / (span / g.num[d]) * (1.0 + 1e-12) + 1e-9;
        if (wr[d] < (int)ceil(reach)) need_win = true;
    }
    const int cap = P.nbr_cap;
    int cnt = 0;
    // the reference's rule for candidate j (pf: its cell-offset record)
    auto accept = [&](int j, const float4& pf) -> bool {
        const uint32_t pk = __float_as_uint(pf.w);
        int cxj, cyj, czj;
        unpack_cell<DIM>(pk, cxj, cyj, czj);
        const float dx = (pfi.x - pf.x) + (float)(cxi - cxj) * csf;
        const float dy = (pfi.y - pf.y) + (float)(cyi - cyj) * csf;
        const float dz = DIM == 3 ? (pfi.z - pf.z) + (float)(czi - czj) * csf : 0.0f;
        const float r2 = (int)(pk | pk_i) < 0 ? lof : dx * dx + dy * dy + dz * dz;
        if (!(r2 < hif)) return false;
        if (!(r2 < lof) && exact_dr<DIM>(qi, D.posh[j], cut, lo2, hi2) < 0.0) return false;
        if (need_win) {
            const uint32_t rcj = D.refcell[j];
            bool inwin = true;
            for (int d = 0; d < DIM; ++d) inwin &= abs(ref_axis(rcj, d) - rc[d]) <= wr[d];
            if (!inwin) return false;
        }
        return true;
    };
    auto take = [&](int j, const float4& pf) {
        if (!accept(j, pf)) return;
        if (cnt < cap) D.nbr[vlist_at(s, cnt, cap)] = j;
        ++cnt;
        visit(j, pf);
    };
    const int nc = D.vcnt[s];
    if (LPT > 1 && nc <= P.verlet_cap) {
        const int* __restrict__ nl = D.vnbr + vlist_at(s, 0, P.verlet_cap);
        const int gbase = (int)(threadIdx.x & 63) - sub;
        const uint64_t below = (1ull << sub) - 1ull;
        for (int b = 0; b < nc; b += LPT) {
            const int k = b + sub;
            int j = 0;
            bool ok = false;
            if (k < nc) {
                j = __builtin_nontemporal_load(nl + (int64_t)k * kVlistLanes);
                ok = accept(j, D.posf[(uint32_t)j]);
            }
            const uint64_t m = (__ballot(ok) >> gbase) & ((1ull << LPT) - 1ull);
            if (ok) {
                const int pos = cnt + __popcll(m & below);
                if (pos < cap) D.nbr[vlist_at(s, pos, cap)] = j;
            }
            cnt += __popcll(m);
        }
        return cnt;
    }
    if (LPT > 1 && sub != 0) return 0;
    if (nc > P.verlet_cap) {
        const uint32_t key = D.key[s];
        const int cx = (int)(key % (uint32_t)P.nc[0]);
        const int rest = (int)(key / (uint32_t)P.nc[0]);
        const int cy = rest % P.nc[1];
        const int cz = rest / P.nc[1];
        const int W = P.win;
        const int zlo = DIM == 3 ? max(cz - W, 0) : 0, zhi = DIM == 3 ? min(cz + W, P.nc[2] - 1) : 0;
        const int ylo = max(cy - W, 0), yhi = min(cy + W, P.nc[1] - 1);
        const int xlo = max(cx - W, 0), xhi = min(cx + W, P.nc[0] - 1);
        for (int z = zlo; z <= zhi; ++z)
            for (int y = ylo; y <= yhi; ++y) {
                const int64_t row = (int64_t)P.nc[0] * ((int64_t)y + (int64_t)P.nc[1] * z);
                const int ke = D.start[row + xhi + 1];
                for (int k = D.start[row + xlo]; k < ke; ++k) {
                    const int j = __float_as_int(D.cpos[k].w);
                    if (j != s) take(j, D.posf[j]);
                }
            }
    } else {
        // the record of candidate k + 1 is in flight while candidate k is tested
        const int* __restrict__ nl = D.vnbr + vlist_at(s, 0, P.verlet_cap);
        int j = nc > 0 ? __builtin_nontemporal_load(nl) : s;
        float4 pj = D.posf[(uint32_t)j];
        for (int k = 0; k < nc; ++k) {
            const int jn = k + 1 < nc ? __builtin_nontemporal_load(nl + (int64_t)(k + 1) * kVlistLanes) : j;
            const float4 pn = D.posf[(uint32_t)jn];
            take(j, pj);
            j = jn;
            pj = pn;
        }
    }
    return cnt;
}

// The count of a target's exact list and its checks: capacity (the step's
// sums would be incomplete) and the largest count.
__device__ __forceinline__ void filter_done(const Params& P, const Dev& D, int s, int cnt) {
    D.ncnt[s] = cnt;
    if (cnt > P.nbr_cap) atomicOr(P.err, kErrPairCap);
    if (cnt > __atomic_load_n(P.maxcnt, __ATOMIC_RELAXED)) atomicMax(P.maxcnt, cnt);
}

// Clears the rebuild gate of the step: every gated kernel ran before the
// launch that calls it (the filter, alone or in the strain pass).
__device__ __forceinline__ void close_gate(const Params& P) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        P.vflag[0] = 0ull;
        P.vflag[1] = 0ull;
    }
}

// The exact lists of this step for every real target (filter_target), when
// no pair pass filters the candidates itself.
// LPT lanes per target (target_slot): 1, or kLanesPerTarget for small runs.
template <int DIM, int LPT>
__global__ __launch_bounds__(kBlock) void k_filter(Params P, Dev D) {
    close_gate(P);
    int sub;
    const int s = target_slot<LPT>(P, nullptr, 0, sub);
    if (s >= P.n) return;
    const int4 mt = D.meta[s];
    if (mt.w >= P.n_real) return;
    if (mt.y == -1) {
        if (sub == 0) D.ncnt[s] = 0;
        return;
    }
    const int cnt = filter_target<DIM, LPT>(P, D, s, D.posh[s], [](int, const float4&) {}, sub);
    if (sub == 0) filter_done(P, D, s, cnt);
}

// The step boundary of the general path: the correct of step n
// (timeIntegration.f90:178-252) fused with the predict + EOS of step n + 1
// (:137-170) for real particles, their FP32 records, the displacement vote,
// and the bounding-box partials of the new positions for the grid of step
// n + 1. Dummies are static: their EOS and records were written by
// k_point_gen and do not change. The reference cells follow once the new grid
// is known (with the keys of the next tail).
//
// One real particle's correct and predict read every field once and write it
// once: the arithmetic of correct_one, predict_one, eos_th and write_rec32 in
// the same order, without reloading what the launch has just stored (each
// such reload waits for the store before it).
template <int DIM>
__device__ __forceinline__ void correct_predict_gen(const Params& P, const Dev& D, int s, int4 mt, bool first,
                                                    double4& q) {
    const double dt = P.dt, h2 = dt / 2;
    const double fac = first ? dt / 2 : dt;
    const double4 vp = D.vprev[s];
    const double4 dv = D.dv[s];
    const double dee = D.de[s];
    const double ep = D.eprev[s];
    double4 v = D.vm[s];
    double4 ds = D.disp[s];
    const double p_old = D.p[s];
    const double c = D.c[s];
    const double rho_mem = P.summation_density ? D.rho[s] : 0.0;
    const bool solid = P.has_solid && mt.x > 100;
    const Material* m = solid ? find_mat(P, mt.x) : nullptr;
    double Sc[DIM * DIM], dS[DIM * DIM];
    if (m) {
        const double* S0 = (first ? D.shear : D.shear_prev) + (int64_t)s * DIM * DIM;
        const double* dS0 = D.dsdt + (int64_t)s * DIM * DIM;
        for (int k = 0; k < DIM * DIM; ++k) {
            Sc[k] = S0[k];
            dS[k] = dS0[k];
        }
    }
    // correct (timeIntegration.f90:178-252)
    double rho = P.summation_density ? rho_mem : add_mul(vp.w, fac, dv.w);
    const bool finite_rho = isfinite(rho);
    double e = add_mul(ep, fac, dee);
    if (e < 0) e = 0;
    v.x = add_mul(vp.x, fac, dv.x);
    v.y = add_mul(vp.y, fac, dv.y);
    if (DIM == 3) v.z = add_mul(vp.z, fac, dv.z);
    if (first) {
        ds.x = dt * v.x; ds.y = dt * v.y; ds.z = DIM == 3 ? dt * v.z : 0.0;
    } else {
        ds.x = add_mul(ds.x, dt, v.x); ds.y = add_mul(ds.y, dt, v.y);
        if (DIM == 3) ds.z = add_mul(ds.z, dt, v.z);
    }
    q.x = add_mul(q.x, dt, v.x);
    q.y = add_mul(q.y, dt, v.y);
    if (DIM == 3) q.z = add_mul(q.z, dt, v.z);
    if (m) {
        for (int k = 0; k < DIM * DIM; ++k) Sc[k] = add_mul(Sc[k], fac, dS[k]);
        if (m->sigma_y > 0) {
            double J2 = 0.0;
            for (int k = 0; k < DIM * DIM; ++k) J2 = add_mul(J2, Sc[k], Sc[k]);
            double sc = fmin(1.0, sqrt(((m->sigma_y * m->sigma_y) / 3) / J2));
            for (int k = 0; k < DIM * DIM; ++k) Sc[k] = Sc[k] * sc;
        }
    }
    if (!isfinite(q.x) || !isfinite(q.y) || !isfinite(q.z) || !isfinite(v.x) || !finite_rho)
        atomicOr(P.err, kErrNonFinite);
    // predict of the next step (:137-170): prev = corrected state, half kick
    D.vprev[s] = make_double4(v.x, v.y, v.z, rho);
    D.eprev[s] = e;
    e = add_mul(e, h2, dee);
    if (e < 0) e = 0;
    if (!P.summation_density) rho = add_mul(rho, h2, dv.w);
    v.x = add_mul(v.x, h2, dv.x);
    v.y = add_mul(v.y, h2, dv.y);
    if (DIM == 3) v.z = add_mul(v.z, h2, dv.z);
    D.vm[s] = v;
    D.rho[s] = rho;
    D.e[s] = e;
    D.disp[s] = ds;
    D.posh[s] = q;
    if (m) {
        double* S = D.shear + (int64_t)s * DIM * DIM;
        double* Sp = D.shear_prev + (int64_t)s * DIM * DIM;
        double Sn[DIM * DIM];
        for (int k = 0; k < DIM * DIM; ++k) {
            Sp[k] = Sc[k];
            Sn[k] = add_mul(Sc[k], h2, dS[k]);
        }
        if (m->sigma_y > 0) {
            double J2 = 0.0;
            for (int k = 0; k < DIM * DIM; ++k) J2 = add_mul(J2, Sn[k], Sn[k]);
            double sc = fmin(1.0, sqrt(((m->sigma_y * m->sigma_y) / 3) / J2));
            for (int k = 0; k < DIM * DIM; ++k) Sn[k] = Sn[k] * sc;
        }
        for (int k = 0; k < DIM * DIM; ++k) S[k] = Sn[k];
    }
    if (P.pa_sph == 3) D.pold[s] = p_old;
    // EOS and the pair inputs th = (rho, p/rho^2, c, m/rho)
    double4 th;
    if (!P.summation_density) {
        double p = p_old;
        if (!(mt.x > 100 || mt.x < -100)) p = eos_p(find_mat(P, mt.x), rho, p);
        D.p[s] = p;
        th = make_double4(rho, p / (rho * rho), c, v.w / rho);
        D.th[s] = th;
    } else if (P.pair32) {
        th = D.th[s];
    }
    D.posf[s] = cell_offset_record(P, q);
    if (P.pair32) {
        PairRecF rec;
        rec.vm = make_float4((float)v.x, (float)v.y, (float)v.z, (float)th.w);
        rec.th = make_float4((float)th.x, (float)th.y, (float)th.z, (float)q.w);
        D.prf[s] = rec;
    }
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_correct_point_gen(Params P, Dev D, int first) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool moved = false;
    if (s < P.n) {
        const int4 mt = D.meta[s];
        double4 q = D.posh[s];
        if (mt.w < P.n_real) {
            correct_predict_gen<DIM>(P, D, s, mt, first != 0, q);
            moved = moved_half_skin<false>(P, D, s, q);
        }
        mn[0] = mx[0] = q.x; mn[1] = mx[1] = q.y; mn[2] = mx[2] = q.z;
    }
    vote_rebuild(P, moved);
    bbox_block_reduce(mn, mx, P.partial, blockIdx.x);
}

// The first pair pass of a general step with the filter in it: each real
// target's exact list from its candidates (filter_target) and its
// strain-rate sums over the neighbours as they are accepted, so each
// neighbour's cell-offset record is loaded once. The list order, and so
// every sum, is that of k_filter followed by k_strain. One lane per target;
// no kernel gradient correction (its pass comes first when it is on).
template <int DIM, typename R>
__global__ __launch_bounds__(kBlock) void k_strain_filter(Params P, Dev D) {
    close_gate(P);
    const int s = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (s >= P.n) return;
    const int4 mt = D.meta[s];
    const Material* m = find_mat(P, mt.x);
    const double eta = (P.viscosity && m) ? m->eta : 0.0;   // viscosity.f90:14-33
    D.eta[s] = eta;
    R T[DIM * DIM];
    for (int q = 0; q < DIM * DIM; ++q) T[q] = (R)0.0;
    if (mt.w < P.n_real) {
        if (mt.y == -1) {
            D.ncnt[s] = 0;
        } else {
            const double4 qi = D.posh[s];
            const double4 vi = D.vm[s];
            const R f = kernel_factor<DIM, R>(P.skf, (R)qi.w);
            const float4 pfi = D.posf[s];
            const int cnt = filter_target<DIM>(P, D, s, qi, [&](int j, const float4& pfj) {
                strain_pair<DIM, R>(P, D, j, &pfj, qi, vi, pfi, f, nullptr, T);
            });
            filter_done(P, D, s, cnt);
        }
    }
    strain_finish<DIM, R>(P, D, s, mt, eta, T);
}
