// Fluid fast path of libsphgpu (included by sphgpu.hip after the shared
// helpers): a candidate list within cutoff + skin, rebuilt only when some
// particle has moved skin/2 since the last build, and one fused pass that
// applies the reference's exact neighbour rule to the candidates and
// accumulates the pair physics of the accepted ones.
//
// Exactness: a pair with r < cut at step n had r < cut + 2 max|x(n) - x(build)|
// < cut + skin at the build, and the build keeps every candidate with r below
// cut + skin (FP32 test with a bounding margin), so the candidate list holds
// every reference neighbour; the fused pass then accepts exactly those with
// the reference's arithmetic (asymmetricSearch.f90:227-237) and window
// (:209-217).

// A batched run keeps its cell tables only while it rebuilds: a real particle
// whose candidates outgrew its list (it walks its cell window in the fused
// pass) votes for the next step's rebuild, so the run rebuilds every step
// until its lists fit again.
template <bool BATCH>
__device__ __forceinline__ bool overflow_vote(const Params& P, const Dev& D, int s, int4 mt) {
    return BATCH && mt.w < P.n_real && mt.y != -1 && D.vcnt[s] > P.verlet_cap;
}

// Predict + EOS + pair inputs for every particle, the reference cell (for the
// window check), and the largest displacement since the last list build.
template <int DIM, bool MIXED, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_point_fast(Params P0, Dev D, int mode, RunMap M) {
    // mode bit 0: first step (no predict); bit 1: no predict at all. In a
    // batch the first step is per run (its step0).
    int lb;
    const Params& P = block_run<BATCH>(P0, M, blockIdx.x, lb);
    const bool first = BATCH ? P.step0 == 1 : (mode & 1) != 0;
    const int l = lb * kBlock + threadIdx.x;
    const int s = P.s0 + l;
    bool moved = false;
    if (l < P.n) {
        const int4 mt = D.meta[s];
        const bool real = mt.w < P.n_real;
        if (MIXED) {
            if (real && !(mode & 2)) predict_mixed<DIM>(P, D, s, first);
            eos_mixed(P, D, s, mt.x);
        } else {
            if (real && !(mode & 2)) predict_one<DIM>(P, D, s, mt, first);
            eos_th(P, D, s, mt);
        }
        const double4 q = D.posh[s];
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
        moved = moved_half_skin<MIXED>(P, D, s, q) || overflow_vote<BATCH>(P, D, s, mt);
    }
    // one atomic per wave, and only when some lane crossed the threshold
    vote_rebuild(P, moved);
}

// Candidate list: every j within cut_i + skin of i (superset is harmless), from
// the +-win sweep over the cell-ordered FP32 positions. Also records the
// build positions of all particles.
template <int DIM, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_verlet_build(Params P0, Dev D, Gate G0, int count_only,
                                                         RunMap RM) {
    int lb;
    const Params& P = block_run<BATCH>(P0, RM, xcd_block(blockIdx.x, gridDim.x), lb);
    const Gate G = BATCH ? Gate{P.vflag} : G0;
    const int l = lb * kBlock + threadIdx.x;
    const int s = P.s0 + l;
    if (!gate_open(G)) return;   // uniform over the block (one run, one gate)
    // every lane stays to the wave reduction at the end; only lanes of
    // particles touch memory
    bool active = false;
    double4 qi = make_double4(0.0, 0.0, 0.0, 0.0);
    if (l < P.n) {
        qi = D.posh[s];
        if (P.mixed) D.xbf[s] = make_float4((float)qi.x, (float)qi.y, (float)qi.z, 0.0f);
        else D.xb[s] = qi;
        const int4 mt = D.meta[s];
        active = mt.w < P.n_real && mt.y != -1;
        if (mt.w < P.n_real && mt.y == -1) D.vcnt[s] = 0;
    }
    const int cap = count_only ? 0 : P.verlet_cap;
    int cnt = 0;
    if (active) {
        // an evolving h (general path) may grow up to the planned limit
        // before the next build
        const double cutv = (P.sle ? P.hmax : qi.w) * (double)P.scale_k + P.skin;
        const RefGrid g = *P.grid;
        const double u = 5.9604644775390625e-8;
        const double e = g.pos_err + u * cutv;
        const double M = 2.0 * 1.7320508075688772 * cutv * e + 3.0 * e * e + 3.0 * u * cutv * cutv;
        const float lim = (float)(cutv * cutv + 2.0 * M + 1e-6 * cutv * cutv);
        const float xf = (float)(qi.x - P.org[0]), yf = (float)(qi.y - P.org[1]), zf = (float)(qi.z - P.org[2]);
        const uint32_t key = D.key[s] - (uint32_t)P.c0;   // cell within the run's grid
        const int* __restrict__ start = D.start + P.c0;
        const int cx = (int)(key % (uint32_t)P.nc[0]);
        const int rest = (int)(key / (uint32_t)P.nc[0]);
        const int cy = rest % P.nc[1];
        const int cz = rest / P.nc[1];
        const int W = P.win;
        const int zlo = DIM == 3 ? max(cz - W, 0) : 0, zhi = DIM == 3 ? min(cz + W, P.nc[2] - 1) : 0;
        const int ylo = max(cy - W, 0), yhi = min(cy + W, P.nc[1] - 1);
        const int xlo = max(cx - W, 0), xhi = min(cx + W, P.nc[0] - 1);
        auto test = [&](const float4 cj) {
            const float dx = xf - cj.x, dy = yf - cj.y, dz = DIM == 3 ? zf - cj.z : 0.0f;
            const float r2 = dx * dx + dy * dy + dz * dz;
            const int j = __float_as_int(cj.w);
            if (r2 < lim && j != s) {
                if (cnt < cap) D.vnbr[vlist_at(s, cnt, cap)] = j;
                ++cnt;
            }
        };
        int z = zlo, y = ylo;
        int64_t row = (int64_t)P.nc[0] * ((int64_t)y + (int64_t)P.nc[1] * z);
        int kb = start[row + xlo], ke = start[row + xhi + 1];
        for (;;) {
            int ny = y + 1, nz = z;
            if (ny > yhi) { ny = ylo; ++nz; }
            const bool more = nz <= zhi;
            int nkb = 0, nke = 0;
            if (more) {
                const int64_t nrow = (int64_t)P.nc[0] * ((int64_t)ny + (int64_t)P.nc[1] * nz);
                nkb = start[nrow + xlo];
                nke = start[nrow + xhi + 1];
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
        D.vcnt[s] = cnt;
    }
    // one atomic per wave for the largest count (same-address atomics from
    // every lane serialise in L2)
    int wmax = cnt;
    for (int o = 32; o > 0; o >>= 1) wmax = max(wmax, __shfl_xor(wmax, o));
    if ((threadIdx.x & 63) == 0 && wmax > __atomic_load_n(P.maxcnt, __ATOMIC_RELAXED)) atomicMax(P.maxcnt, wmax);
}

// Exact neighbour test of one candidate in FP64, the reference's arithmetic:
// squared distance summed in axis order, correctly rounded, no contraction;
// accept sqrt(dr) < cut. Returns dr (or -1 if rejected).
template <int DIM>
__device__ __forceinline__ double exact_dr(const double4& qi, const double4& qj, double cut,
                                           double lo2, double hi2) {
    const double dx0 = rn_sub(qi.x, qj.x);
    const double dx1 = rn_sub(qi.y, qj.y);
    double dr = rn_add(rn_mul(dx0, dx0), rn_mul(dx1, dx1));
    if (DIM == 3) {
        const double dx2 = rn_sub(qi.z, qj.z);
        dr = rn_add(dr, rn_mul(dx2, dx2));
    }
    if (dr < lo2) return dr;
    if (!(dr < hi2)) return -1.0;
    return (__dsqrt_rn(dr) < cut) ? dr : -1.0;
}

// Fused pass over the candidate list: exact neighbour test + window, then the
// same physics as k_force_fluid for the accepted pairs.
template <int DIM, bool MIXED, bool PAIR_EXT, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_force_fused(Params P0, Dev D, RunMap RM) {
    using R = typename std::conditional<MIXED, float, double>::type;
    using R4 = typename Vec4<R>::T;
    int lb;
    const Params& P = block_run<BATCH>(P0, RM, xcd_block(blockIdx.x, gridDim.x), lb);
    if (threadIdx.x == 0) {
        // every gated kernel of this step has run: clear the gates for the next
        if (lb == 0) {
            P.vflag[0] = 0ull;
            P.vflag[1] = 0ull;
        }
        if (BATCH && blockIdx.x == 0) {
            P.vany[0] = 0ull;
            P.vany[1] = 0ull;
        }
    }
    const int l = lb * kBlock + threadIdx.x;
    const int s = P.s0 + l;
    if (l >= P.n) return;
    const int4 mt = D.meta[s];
    if (mt.w >= P.n_real) return;
    const int ti = mt.x;
    const double4 qi = D.posh[s];
    const double cut = qi.w * (double)P.scale_k;
    const double cut2 = cut * cut;
    const double lo2 = cut2 * (1.0 - 1e-12), hi2 = cut2 * (1.0 + 1e-12);
    R4 vi, thi;
    R pix = 0, piy = 0, piz = 0;
    int cxi = 0, cyi = 0, czi = 0;
    uint32_t pk_i = 0;
    float lof = 0.0f, hif = 0.0f;
    if (MIXED) {
        vi = *(const R4*)&D.prf[s].vm;
        thi = *(const R4*)&D.prf[s].th;
        const float4 pf = D.posf[s];
        pix = pf.x; piy = pf.y; piz = pf.z;
        pk_i = __float_as_uint(pf.w);
        unpack_cell<DIM>(pk_i, cxi, cyi, czi);
        // FP32 r^2 from cell offsets is within M of the FP64 one
        const double u = 5.9604644775390625e-8;
        const double e = 2.0 * u * P.cs + u * cut;
        const double M = 2.0 * 1.7320508075688772 * cut * e + 3.0 * e * e + 3.0 * u * cut2;
        lof = (float)(cut2 - 2.0 * M - 1e-6 * cut2);
        hif = (float)(cut2 + 2.0 * M + 1e-6 * cut2);
    } else {
        vi = *(const R4*)&D.vm[s];
        thi = *(const R4*)&D.th[s];
    }
    const R hi = (R)qi.w;
    // reference window (only when it can exclude a pair of this particle)
    const RefGrid g = *P.grid;
    int wr[3] = {0, 0, 0}, rc[3] = {0, 0, 0};
    bool need_win = false;
    const uint32_t rci = g.need_win ? D.refcell[s] : 0u;
    for (int d = 0; d < DIM && g.need_win; ++d) {
        const double span = g.maxc[d] - g.minc[d];
        wr[d] = (int)(qi.w / span * g.num[d]) + 1;
        rc[d] = ref_axis(rci, d);
        const double reach = cut / (span / g.num[d]) * (1.0 + 1e-12) + 1e-9;
        if (wr[d] < (int)ceil(reach)) need_win = true;
    }
    const R csf = (R)P.cs;
    const R inv_h = (R)1 / hi;
    const R fac = DIM == 2 ? (R)(15.0 / (7.0 * kPi)) * inv_h * inv_h
                           : (R)(3.0 / (2.0 * kPi)) * inv_h * inv_h * inv_h;
    const R g1 = fac * inv_h * inv_h;
    const R g2 = (R)-0.5 * fac * inv_h;
    const R alpha = (R)P.av_alpha, beta = (R)P.av_beta, psi2 = (R)(P.av_psi * P.av_psi);
    R ax = 0, ay = 0, az = 0, de_in = 0, de_av = 0, sv = 0, sx = 0;
    int acc_n = 0;
    // a particle with more candidates than its list holds walks its cell
    // window of the last sort instead: that window holds every candidate the
    // list would (the sort is at or after the list's build), and the exact
    // test keeps only the reference's pairs. A batch's shared cell tables
    // hold a run only in the steps it rebuilds, so a batched run with such a
    // particle votes for a rebuild every step (overflow_vote).
    const int nc = (mt.y == -1) ? 0 : D.vcnt[s];
    const bool sweep = nc > P.verlet_cap;
    const int n = sweep ? 0 : min(nc, P.verlet_cap);
    const int* __restrict__ nl = D.vnbr + vlist_at(s, 0, P.verlet_cap);
    // the mixed pair sums in fewer operations (the same quantities, FP32
    // rounding): grad W = G dx, so dv . grad W = G (dx . dv) and
    // dx . grad W = G r^2; the energy rate 1/2 sum m_j (aux + Pi) dv . grad W
    // is 1/2 sum cf dv . grad W; and with A = (h_i + h_j) dx . dv and
    // B = r^2 + psi^2 hbar^2, Pi = (beta A / 2 - alpha (c_i + c_j) B / 2) A /
    // (B^2 (rho_i + rho_j)): one reciprocal instead of two
    const R beta_h = (R)0.5 * beta, alpha_h = (R)0.5 * alpha, psi2_q = (R)0.25 * psi2;
    // physics of one accepted pair, both precisions
    auto sum_pair = [&](int j, R dx, R dy, R dz, R r2, R r, R G, const R4& vj, const R4& thj, R hj, R Vj,
                        R mj) {
        const R dvx = vi.x - vj.x, dvy = vi.y - vj.y, dvz = DIM == 3 ? vi.z - vj.z : (R)0;
        const R xv = dx * dvx + dy * dvy + dz * dvz;
        const R aux = thi.y + thj.y;
        if constexpr (MIXED) {
            const R dvg = G * xv;
            sv += Vj * dvg;
            sx += Vj * (G * r2);
            // evaluated for every pair and selected: a branch here lets the
            // compiler split the neighbour record load and wait twice
            const R hs = hi + hj;
            const R A = hs * xv;
            const R B = psi2_q * hs * hs + r2;
            const R num = beta_h * A - alpha_h * (thi.z + thj.z) * B;
            const R piv = num * A * fast_rcp(B * B * (thi.x + thj.x));
            const R Pi = xv < (R)0 ? piv : (R)0;
            const R cf = mj * (aux + Pi);
            de_in += cf * dvg;
            const R cfg = cf * G;
            ax -= cfg * dx;
            ay -= cfg * dy;
            az -= cfg * dz;
        } else {
            ++acc_n;
            const R gx = G * dx, gy = G * dy, gz = G * dz;
            const R dvg = dvx * gx + dvy * gy + dvz * gz;
            const R dxg = dx * gx + dy * gy + dz * gz;
            sv += Vj * dvg;
            sx += Vj * dxg;
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
        }
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
    };
    auto in_window = [&](int j) {
        if (!need_win) return true;
        const uint32_t rcj = D.refcell[j];
        bool inwin = true;
        for (int d = 0; d < DIM; ++d) inwin &= abs(ref_axis(rcj, d) - rc[d]) <= wr[d];
        return inwin;
    };
    // every particle but s in the +-win cell window around s's cell of the
    // last sort, in cell order (the order of the list build)
    auto for_window = [&](auto&& f) {
        const uint32_t key = D.key[s] - (uint32_t)P.c0;
        const int* __restrict__ start = D.start + P.c0;
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
                const int ke = start[row + xhi + 1];
                for (int k = start[row + xlo]; k < ke; ++k) {
                    const int j = __float_as_int(D.cpos[k].w);
                    if (j != s) f(j);
                }
            }
    };
    if constexpr (MIXED) {
        // the window test |c_j - c_i| <= w of each axis as one unsigned
        // compare, c_j - c_i + w <= 2w (cell fields are 10 bits); a lane whose
        // own window cannot exclude a pair takes a bias that passes every cell
        uint32_t wb[3], wl[3];
        for (int d = 0; d < 3; ++d) {
            wb[d] = need_win ? (uint32_t)(wr[d] - rc[d]) : 2048u;
            wl[d] = need_win ? 2u * (uint32_t)wr[d] : 4096u;
        }
        auto in_window_rc = [&](uint32_t rcj) {
            bool inwin = true;
            for (int d = 0; d < DIM; ++d) inwin &= (uint32_t)ref_axis(rcj, d) + wb[d] <= wl[d];
            return inwin;
        };
        // the candidate loop in two forms, chosen once per launch: with the
        // window test, each candidate's window cell read with its records (a
        // load inside the window branch would hold the wave until every record
        // load in flight had returned), or, where no particle's window can
        // exclude a pair, with neither
        auto run = [&](auto win_tag) {
            constexpr bool WIN = decltype(win_tag)::value;
            auto candidate = [&](int j, const float4& pf, const float4& vj, const float4& thj, uint32_t rcj) {
                const uint32_t pk = __float_as_uint(pf.w);
                int cxj, cyj, czj;
                unpack_cell<DIM>(pk, cxj, cyj, czj);
                const R dx = (pix - pf.x) + (R)(cxi - cxj) * csf;
                const R dy = (piy - pf.y) + (R)(cyi - cyj) * csf;
                const R dz = DIM == 3 ? (piz - pf.z) + (R)(czi - czj) * csf : (R)0;
                const R r2 = dx * dx + dy * dy + dz * dz;
                // acceptance in bitwise logic (one mask, no nested regions); the
                // FP64 test runs only where the FP32 distance is within its error
                // bound of the cutoff or a particle is off the grid (bit 31 of the
                // cell word): such a pair is tested at r2 = lof, inside the band
                const R r2t = (int)(pk | pk_i) < 0 ? lof : r2;
                bool ok = r2t < hif;
                if (ok & !(r2t < lof)) ok = exact_dr<DIM>(qi, D.posh[j], cut, lo2, hi2) >= 0.0;
                if constexpr (WIN) ok = ok & in_window_rc(rcj);
                // every candidate runs the pair arithmetic and a rejected one adds
                // zeros (G = 0): no divergent branch around the sums. vm.w = m / rho;
                // one rsq gives r and 1 / r; both kernel branches are cheap, so
                // select instead of diverging
                const R rinv = fast_rsq(r2);
                const R r = r2 * rinv;
                const R q = r * inv_h;
                const R t = (R)2 - q;
                const R Gn = g1 * ((R)-2 + (R)1.5 * q);
                const R Gf = g2 * t * t * rinv;
                acc_n += ok ? 1 : 0;
                sum_pair(j, dx, dy, dz, r2, r, ok ? (q < (R)1 ? Gn : Gf) : (R)0, vj, thj, thj.w, vj.w,
                         vj.w * thj.x);
            };
            // software pipeline two candidates deep: the records of candidates
            // k + 1 and k + 2 are in flight while candidate k is tested and
            // summed (loaded whether or not they turn out to be neighbours).
            // Three record sets take turns in a three-candidate body, so no set
            // is copied (a copy would wait for its loads). The candidate list is
            // read once per pass: nontemporal loads keep it from evicting
            // neighbour records from L1/L2. Unsigned slot indices make each
            // record address one shift-add.
            constexpr int ldn = kVlistLanes;
            struct Rec {
                int j;
                float4 pf, vm, th;
                uint32_t rc;
            };
            auto load = [&](Rec& r, int j) {
                r.j = j;
                r.pf = D.posf[(uint32_t)j];
                r.vm = D.prf[(uint32_t)j].vm;
                r.th = D.prf[(uint32_t)j].th;
                r.rc = WIN ? D.refcell[(uint32_t)j] : 0u;
            };
            // an index past the end reads this particle's own records (never
            // summed); it depends on no load in flight
            auto index = [&](int k) { return k < n ? __builtin_nontemporal_load(nl + (int64_t)k * ldn) : s; };
            const int i0 = index(0), i1 = index(1);
            int i2 = index(2), i3 = index(3), i4 = index(4);
            Rec A, B, C;
            load(A, i0);
            load(B, i1);
            for (int k = 0; k < n; k += 3) {
                const int i5 = index(k + 5), i6 = index(k + 6), i7 = index(k + 7);
                load(C, i2);
                candidate(A.j, A.pf, A.vm, A.th, A.rc);
                load(A, i3);
                if (k + 1 < n) candidate(B.j, B.pf, B.vm, B.th, B.rc);
                load(B, i4);
                if (k + 2 < n) candidate(C.j, C.pf, C.vm, C.th, C.rc);
                i2 = i5;
                i3 = i6;
                i4 = i7;
            }
            if (sweep)
                for_window([&](int j) {
                    candidate(j, D.posf[(uint32_t)j], D.prf[(uint32_t)j].vm, D.prf[(uint32_t)j].th,
                              WIN ? D.refcell[(uint32_t)j] : 0u);
                });
        };
        if (g.need_win)
            run(std::true_type{});
        else
            run(std::false_type{});
    } else {
        auto candidate64 = [&](int j) {
            const double4 qj = D.posh[j];
            const double dr = exact_dr<DIM>(qi, qj, cut, lo2, hi2);
            if (dr < 0.0) return;
            if (!in_window(j)) return;
            const R4 vj = *(const R4*)&D.vm[j];
            const R4 thj = *(const R4*)&D.th[j];
            const R r = fast_sqrt((R)dr);
            const R q = r * inv_h;
            const R t = (R)2 - q;
            const R G = q < (R)1 ? g1 * ((R)-2 + (R)1.5 * q) : g2 * t * t * fast_rcp(r);
            sum_pair(j, (R)(qi.x - qj.x), (R)(qi.y - qj.y), DIM == 3 ? (R)(qi.z - qj.z) : (R)0, (R)dr, r, G,
                     vj, thj, (R)qj.w, thj.w, vj.w);
        };
        int jn = n > 0 ? __builtin_nontemporal_load(nl) : 0;
        for (int k = 0; k < n; ++k) {
            const int j = jn;
            if (k + 1 < n) jn = __builtin_nontemporal_load(nl + (int64_t)(k + 1) * kVlistLanes);
            candidate64(j);
        }
        if (sweep) for_window(candidate64);
    }
    if (P.gravity && !(P.gravity_mode == SPHGPU_GRAVITY_SKIP_SOLID && ti > 100)) {
        if (DIM == 2) ay += (R)P.g;
        else az += (R)P.g;
    }
    if (mt.z == 1) ax = ay = az = 0;
    const R drho = thi.x * sv;
    const R de = (R)0.5 * de_in + de_av;   // mixed: de_in holds both terms, de_av is 0
    if (MIXED) {
        D.dvf[s] = make_float4((float)ax, (float)ay, (float)az, (float)drho);
        D.def[s] = (float)de;
    } else {
        D.dv[s] = make_double4(ax, ay, az, drho);
        D.de[s] = de;
    }
    if (MIXED) D.divf[s] = make_float2(-(float)sx, -(float)sv);
    else D.div[s] = make_double2(-(double)sx, -(double)sv);
    D.ecnt[s] = acc_n;
}

// Leap-frog correct of step n-1 (timeIntegration.f90:178-252) fused with the
// predict + EOS of step n (:137-170, internalForce.f90:103-134) and the
// displacement check, for real particles; bounding-box partials of the new
// positions for the grid of step n. Dummies are static: their pair inputs
// were written by k_point_fast and do not change.
template <int DIM, bool MIXED, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_correct_predict(Params P0, Dev D, int first_arg, RunMap M,
                                                            int kstep) {
    int lb;
    const Params& P = block_run<BATCH>(P0, M, blockIdx.x, lb);
    const bool first = BATCH ? (P.step0 + kstep == 1) : first_arg != 0;
    const int l = lb * kBlock + threadIdx.x;
    const int s = P.s0 + l;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool moved = false;
    if (l < P.n) {
        const int4 mt = D.meta[s];
        double4 q = D.posh[s];
        if (mt.w < P.n_real) {
            const double dt = P.dt, h2 = dt / 2;
            const double fac = first ? h2 : dt;
            const double4 vp = D.vprev[s];
            double dvx, dvy, dvz, dee;
            bool finite_rho;
            // mixed: vprev.w holds the mass and the FP64 predicted velocity is
            // not stored (the pair pass reads its FP32 copy; the closing
            // correct writes the FP64 state)
            double4 v = MIXED ? make_double4(0.0, 0.0, 0.0, vp.w) : D.vm[s];
            double rho_c = 0.0;
            float dr_pred = 0.0f;
            if (MIXED) {
                const float4 dv = D.dvf[s];
                dvx = dv.x; dvy = dv.y; dvz = dv.z; dee = D.def[s];
                float dr = add_mul(D.drho_prev[s], (float)fac, dv.w);   // correct
                finite_rho = isfinite(dr);
                D.drho_prev[s] = dr;
                dr_pred = add_mul(dr, (float)h2, dv.w);               // predict of the next step
                D.drho[s] = dr_pred;
            } else {
                const double4 dv = D.dv[s];
                dvx = dv.x; dvy = dv.y; dvz = dv.z; dee = D.de[s];
                rho_c = add_mul(vp.w, fac, dv.w);                    // correct
                finite_rho = isfinite(rho_c);
                D.rho[s] = add_mul(rho_c, h2, dv.w);                 // predict of the next step
            }
            energy_correct_predict<MIXED>(D, s, fac, h2, dee);
            v.x = add_mul(vp.x, fac, dvx);
            v.y = add_mul(vp.y, fac, dvy);
            if (DIM == 3) v.z = add_mul(vp.z, fac, dvz);
            add_disp<DIM, MIXED>(D, s, first, dt, v);
            q.x = add_mul(q.x, dt, v.x);
            q.y = add_mul(q.y, dt, v.y);
            if (DIM == 3) q.z = add_mul(q.z, dt, v.z);
            D.posh[s] = q;
            // predict of the next step: prev = corrected state, then half kick
            D.vprev[s] = make_double4(v.x, v.y, v.z, MIXED ? vp.w : rho_c);
            v.x = add_mul(v.x, h2, dvx);
            v.y = add_mul(v.y, h2, dvy);
            if (DIM == 3) v.z = add_mul(v.z, h2, dvz);
            if (!MIXED) D.vm[s] = v;
            if (!isfinite(q.x) || !isfinite(q.y) || !isfinite(q.z) || !isfinite(v.x) || !finite_rho)
                atomicOr(P.err, kErrNonFinite);
            if (MIXED) eos_mixed_vals(P, D, s, mt.x, dr_pred, v, q);
            else eos_th(P, D, s, mt);
            moved = moved_half_skin<MIXED>(P, D, s, q) || overflow_vote<BATCH>(P, D, s, mt);
        }
        mn[0] = mx[0] = q.x; mn[1] = mx[1] = q.y; mn[2] = mx[2] = q.z;
    }
    vote_rebuild(P, moved);
    bbox_block_reduce(mn, mx, P.partial, lb);
}

// Reference cells of all particles, needed only when the reference window can
// exclude pairs (grid flag); otherwise the launch returns at once.
template <int DIM, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_refcell(Params P0, Dev D, RunMap M) {
    if (!BATCH && !P0.grid->need_win) return;
    const int lim = BATCH ? M.nslot : P0.n;
    for (int s = blockIdx.x * kBlock + threadIdx.x; s < lim; s += gridDim.x * kBlock) {
        const Params* Pp = slot_run<BATCH>(P0, M, s);
        if (!Pp || !Pp->grid->need_win) continue;
        const Params& P = *Pp;
        const RefGrid g = *P.grid;
        const int4 mt = D.meta[s];
        if (mt.y == -1) continue;
        const double4 q = D.posh[s];
        bool bad = false;
        const double xs[3] = {q.x, q.y, q.z};
        int rc[3] = {1, 1, 1};
        for (int d = 0; d < DIM; ++d) rc[d] = ref_locate(xs[d], g.minc[d], g.maxc[d], g.num[d], &bad);
        if (bad) atomicOr(P.err, (isfinite(q.x) && isfinite(q.y) && isfinite(q.z)) ? kErrOutOfRange
                                                                                  : kErrNonFinite);
        D.refcell[s] = pack_ref(rc[0], rc[1], rc[2]);
    }
}

__global__ void k_set_flag(unsigned long long* flag) {  // force the next rebuild
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        flag[0] = 1ull;
        flag[1] = 0ull;
    }
}
