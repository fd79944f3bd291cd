// qtp_host_setup.hpp -- host side of the gfx950 qTIP4P/F step: context
// set-up (the reference's derived parameters, k-groups and the matrix-core
// k-space layout, the normal-mode matrix, allocation).
// Included once, in order, by qtip4pf.hip (one translation unit).
#pragma once

// ---------------------------------------------------------------------------
// set-up
// ---------------------------------------------------------------------------
static void derive(qtp_ctx *ctx)
{
    const qtp_params &P = ctx->prm;
    qtp_derived &d = ctx->drv;
    d.n_mols = P.n_mols;
    d.n_atoms = 3 * P.n_mols;
    d.n_dof = 3 * d.n_atoms;
    d.n_beads = P.n_beads;
    for (int k = 0; k < 3; ++k) d.boxl[k] = P.boxl[k];
    d.vol = P.boxl[0] * P.boxl[1] * P.boxl[2];
    const double rho_ox = P.n_mols / d.vol;
    // qtip4pfparams.h:41, 109-110 (the cutoff optionally shorter)
    d.lj_rcut = P.lj_rcut > 0.0 ? P.lj_rcut : RC.r_cut;
    d.r_skin = 1.30 * d.lj_rcut;
    d.list_size = P.list_size > 0
                      ? P.list_size
                      : (int)std::round(4.0 * d.r_skin * d.r_skin * d.r_skin * P.n_mols * rho_ox);
    // water.cpp:76-83, 138-157
    const double T_red = RC.K2au * P.T_K;
    d.beta = 1.0 / T_red;
    d.piovern = RC.pi / P.n_beads;
    d.beta_n = d.beta / P.n_beads;
    d.omega_p = 1.0 / (d.beta_n * 1.0);
    d.omega_psq = d.omega_p * d.omega_p;
    d.sd_ox = std::sqrt(RC.mo / d.beta_n);
    d.sd_h = std::sqrt(RC.mh / d.beta_n);
    d.dt = P.dtfs * RC.fs2au;
    d.halfdt = 0.5 * d.dt;
    d.epsilonsc = P.epsilon_sc;
    d.acc = std::sqrt(-std::log(P.eps_ewald));
    d.rho = d.n_atoms / d.vol;
    const double min_xy = std::min(P.boxl[0], P.boxl[1]);
    d.min_box = std::min(min_xy, P.boxl[2]);
    d.kwald_prefactor = 2.0 * RC.pi / d.vol;
    const double max_xy = std::max(P.boxl[0], P.boxl[1]);
    d.max_box = std::max(max_xy, P.boxl[2]);
    // forces.cpp:574-584
    if (P.ewald_kappa > 0.0) {
        d.kappa = P.ewald_kappa;
        d.rcut = P.ewald_rcut;
        d.k_cut = P.ewald_kcut;
        d.k_max = P.ewald_kmax;
    } else {
        double kappa = 2.6 * std::pow(d.rho / d.vol, 1.0 / 6.0);
        double rcut = d.acc / kappa;
        if (rcut > 0.5 * d.min_box) {
            rcut = 0.5 * d.min_box;
            kappa = d.acc / rcut;
        }
        d.kappa = kappa;
        d.rcut = rcut;
        d.k_cut = 2.0 * d.acc * kappa;
        d.k_max = (int)(d.k_cut * d.max_box / (2.0 * RC.pi));
    }
}

// k-vector groups in the reference's inclusion rules (forces.cpp:251-445):
// kx-only: 0 < k^2 < k_cut^2; (kx,0,kz): kz loop breaks at k^2 > k_cut^2;
// (kx,ky,0): ky loop breaks likewise; (kx,ky,kz) inside an unbroken ky.
// Coefficient c = the reference's exp_fac (kx-only: kwald_prefactor * exp_fac).
static void build_groups(qtp_ctx *ctx, std::vector<int4> &g, std::vector<double> &cf,
                         std::vector<double4> &rk)
{
    const qtp_derived &d = ctx->drv;
    const int K = d.k_max + 1;
    const double kc2 = d.k_cut * d.k_cut;
    const double q4 = 0.25 / (d.kappa * d.kappa);
    double bk[3];
    for (int k = 0; k < 3; ++k) bk[k] = 2.0 * RC.pi / d.boxl[k];
    const double kp = d.kwald_prefactor;
    for (int kx = 0; kx < K; ++kx) {
        const double r0 = bk[0] * kx;
        double ksq = r0 * r0;
        if (ksq > 0.0 && ksq < kc2) {
            double ef0 = std::exp(-ksq * q4) / ksq;
            if (kx != 0) ef0 *= 2.0;
            g.push_back(make_int4(kx, 0, 0, 1));
            cf.push_back(kp * ef0);
            rk.push_back(make_double4(r0, 0.0, 0.0, 0.0));
        }
        for (int kz = 1; kz < K; ++kz) {
            const double r2 = bk[2] * kz;
            ksq = r0 * r0 + r2 * r2;
            if (ksq > kc2) break;
            double ef = kp * std::exp(-ksq * q4) / ksq;
            if (kx != 0) ef *= 2.0;
            g.push_back(make_int4(kx, 0, kz, 3));
            cf.push_back(ef);
            rk.push_back(make_double4(r0, 0.0, r2, 0.0));
        }
        for (int ky = 1; ky < K; ++ky) {
            const double r1 = bk[1] * ky;
            ksq = r0 * r0 + r1 * r1;
            if (ksq > kc2) break;
            double ef = kp * std::exp(-ksq * q4) / ksq;
            if (kx != 0) ef *= 2.0;
            g.push_back(make_int4(kx, ky, 0, 5));
            cf.push_back(ef);
            rk.push_back(make_double4(r0, r1, 0.0, 0.0));
            for (int kz = 1; kz < K; ++kz) {
                const double r2 = bk[2] * kz;
                ksq = r0 * r0 + r1 * r1 + r2 * r2;
                if (ksq > kc2) break;
                double ef3 = kp * std::exp(-ksq * q4) / ksq;
                if (kx != 0) ef3 *= 2.0;
                g.push_back(make_int4(kx, ky, kz, 15));
                cf.push_back(ef3);
                rk.push_back(make_double4(r0, r1, r2, 0.0));
            }
        }
    }
}

// Matrix layout of k_kforce_mf.  The groups come in columns of fixed
// (kx, ky) and rising kz (build_groups' order); the columns are sorted by
// their largest kz (descending, stable) and paired into 16-row tiles, each
// taking ceil((2 kz_max + 2) / 4) k-steps of 4 (tiles: in that sorted
// order; mf_tile_order arranges them for a block of W waves).  Returns
// false when a (kx, ky) pair appears in two runs or a column's kz are not
// consecutive (then the vector kernels are used).
// k_ksum_mf: kcol per sorted column (kx | ky << 16, first group, first kz,
// last kz), padded to a multiple of four with empty columns; units (tile of
// four columns, n-tile) for every n = 2 kz + 0/1 that a tile's columns reach.
static bool build_mf_layout(const std::vector<int4> &g, std::vector<int> &gcol,
                            std::vector<int4> &tiles, int &Mp, int &nks, std::vector<int4> &kcol,
                            std::vector<int2> &kunit)
{
    struct Col {
        int kx, ky, kzmax, g0, kz0;
    };
    std::vector<Col> cols;
    std::vector<int> gc(g.size());
    std::map<int, int> seen;
    for (size_t i = 0; i < g.size(); ++i) {
        if (cols.empty() || cols.back().kx != g[i].x || cols.back().ky != g[i].y) {
            if (!seen.emplace(g[i].x | (g[i].y << 16), (int)cols.size()).second) return false;
            cols.push_back({g[i].x, g[i].y, g[i].z, (int)i, g[i].z});
        } else {
            if (g[i].z != cols.back().kzmax + 1) return false;
            cols.back().kzmax = g[i].z;
        }
        gc[i] = (int)cols.size() - 1;
    }
    const int nc = (int)cols.size();
    std::vector<int> ord(nc), rank(nc);
    for (int i = 0; i < nc; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int a, int b) { return cols[a].kzmax > cols[b].kzmax; });
    for (int i = 0; i < nc; ++i) rank[ord[i]] = i;
    gcol.resize(g.size());
    for (size_t i = 0; i < g.size(); ++i) gcol[i] = rank[gc[i]];
    const int ntl = (nc + 1) / 2;
    Mp = 16 * ntl;
    nks = 0;
    tiles.assign(ntl, make_int4(0, 0, 0, 0));
    for (int t = 0; t < ntl; ++t) {
        const Col &c0 = cols[ord[2 * t]];
        int4 x;
        x.x = c0.kx | (c0.ky << 16);
        x.y = 2 * t + 1 < nc ? (cols[ord[2 * t + 1]].kx | (cols[ord[2 * t + 1]].ky << 16)) : 0;
        x.z = (2 * c0.kzmax + 2 + 3) / 4;  // the first column has the larger kz_max
        x.w = 16 * t;
        nks = std::max(nks, x.z);
        tiles[t] = x;
    }
    const int nq = (nc + 3) / 4;
    kcol.assign((size_t)4 * nq, make_int4(0, 0, 1, 0));
    for (int i = 0; i < nc; ++i) {
        const Col &c = cols[ord[i]];
        kcol[i] = make_int4(c.kx | (c.ky << 16), c.g0, c.kz0, c.kzmax);
    }
    kunit.clear();
    for (int t = 0; t < nq; ++t)
        for (int u = 0; u <= (2 * kcol[4 * t].w + 1) / 16; ++u) kunit.push_back(make_int2(t, u));
    return true;
}

// The tile table of k_kforce_mf for blocks of W waves: wave w reads entries
// w, w + W, ...; round q hands the sorted tiles q W ... out to the waves in
// turn, reversed in odd rounds, which evens out the k-steps per wave (empty
// entries: k-steps 0).
static std::vector<int4> mf_tile_order(const std::vector<int4> &tiles, int W)
{
    const int ntl = (int)tiles.size();
    const int rounds = (ntl + W - 1) / W;
    std::vector<int4> rt((size_t)rounds * W, make_int4(0, 0, 0, 0));
    for (int t = 0; t < ntl; ++t) {
        const int q = t / W, p = t - q * W;
        rt[(size_t)q * W + ((q & 1) ? W - 1 - p : p)] = tiles[t];
    }
    return rt;
}

// Orthonormal real normal-mode matrix (FFTW R2HC halfcomplex layout with the
// scalings of fourier.cpp:22-66), entries from long double with exact pi.
static void build_nm(int P, std::vector<double> &C)
{
    C.assign((size_t)P * P, 0.0);
    const long double twopi = 6.283185307179586476925286766559005768L;
    const long double s1 = sqrtl(1.0L / P), s2 = sqrtl(2.0L / P);
    for (int b = 0; b < P; ++b) {
        C[0 * P + b] = (double)s1;
        for (int k = 1; k < P - k; ++k) {
            const long double ang = twopi * (long double)(((long)k * b) % P) / P;
            C[(size_t)k * P + b] = (double)(s2 * cosl(ang));
            C[(size_t)(P - k) * P + b] = (double)(-s2 * sinl(ang));
        }
        if (P % 2 == 0) C[(size_t)(P / 2) * P + b] = (double)(s1 * ((b % 2) ? -1.0L : 1.0L));
    }
}

// Cell grid of the cell-binned list build: cells at least r_skin / S wide
// (relative margin 1e-9, far above the rounding of a cell index), S = 2 when
// every side holds 5 such cells, else S = 1 with 3; false when neither fits,
// or the bins or row bitmaps would not fit in LDS.
static bool cell_grid(double r_skin, const double L[3], int n_mols, CellGrid &cg)
{
    if (n_mols > CELL_MAXMOL) return false;
    for (int S = 2; S >= 1; --S) {
        int nc[3];
        bool ok = true;
        for (int d = 0; d < 3; ++d) {
            nc[d] = (int)std::floor(S * L[d] / (r_skin * (1.0 + 1e-9)));
            ok = ok && nc[d] >= 2 * S + 1;
        }
        if (!ok || (long)nc[0] * nc[1] * nc[2] > CELL_MAXCELL) continue;
        cg.S = S;
        cg.ncell = nc[0] * nc[1] * nc[2];
        for (int d = 0; d < 3; ++d) {
            cg.nc[d] = nc[d];
            cg.sc[d] = nc[d] / L[d];
        }
        return true;
    }
    return false;
}

// Lanes per molecule of k_ljintra (16 or 64) and the energy-partial layout
// that follows from it: LJ/intra blocks, real-space tiles, k-space blocks,
// spline blocks.
static void set_lj_lanes(qtp_ctx *ctx, int lpm)
{
    ctx->lj_lpm = lpm;
    ctx->nmb = (ctx->n_mols + (256 / lpm) - 1) / (256 / lpm);
    ctx->ep_lj = 0;
    ctx->ep_real = ctx->nmb;
    ctx->ep_k = ctx->nmb + ctx->ntile;
    ctx->ep_spl = ctx->nmb + ctx->ntile + ctx->ngr;
    ctx->ep_stride = ctx->ep_spl + ctx->nblk;
}

// whether the next list build takes the cell-binned kernels
static bool use_cells(const qtp_ctx *ctx)
{
    if (!ctx->cells_ok || ctx->list_method == QTP_LIST_ALLPAIRS) return false;
    return ctx->list_method == QTP_LIST_CELLS || ctx->n_mols > NL_FUSED_MAXMOL;
}

// Ewald self energy sum_i (kappa / sqrt(pi)) q_i^2 (forces.cpp:622-624), in
// the reference's per-site order but with an 80-bit accumulator: the FP64
// running sum of thousands of equal terms drifts by up to ~5e-14 relative
// (3e-11 Eh of the 571 Eh at 4096 waters), which V, a difference of such
// terms, would carry at ~1e-11 relative.
static double self_energy(const qtp_ctx *ctx, const double *charge)
{
    const double pref = 2.0 * ctx->drv.kappa / std::sqrt(RC.pi);
    long double es = 0.0L;
    for (int a = 0; a < ctx->n_atoms; ++a) es += (long double)(0.5 * pref * charge[a] * charge[a]);
    return (double)es;
}

static int alloc_all(qtp_ctx *ctx)
{
    const int P = ctx->P, E = std::max(P, 1);
    const long N3 = 3L * ctx->NAP;
    CHK(dalloc(ctx, &ctx->q, (size_t)P * N3));
    CHK(dalloc(ctx, &ctx->p, (size_t)P * N3));
    CHK(dalloc(ctx, &ctx->F, (size_t)P * N3));
    CHK(dalloc(ctx, &ctx->fph, (size_t)P * N3));
    CHK(dalloc(ctx, &ctx->disp, (size_t)P * 3 * ctx->NMP));
    CHK(dalloc(ctx, &ctx->m_atom, ctx->NAP));
    CHK(dalloc(ctx, &ctx->sqm_atom, ctx->NAP));
    CHK(dalloc(ctx, &ctx->charge, ctx->NAP));
    CHK(dalloc(ctx, &ctx->point, (size_t)P * ctx->n_mols));
    CHK(dalloc(ctx, &ctx->list, (size_t)P * ctx->list_size));
    CHK(dalloc(ctx, &ctx->fpoint, (size_t)P * (ctx->n_mols + 1)));
    CHK(dalloc(ctx, &ctx->flist, (size_t)P * ctx->fcap));
    CHK(dalloc(ctx, &ctx->cnt, (size_t)P * 2 * ctx->n_mols));
    if (ctx->cells_ok) {
        CHK(dalloc(ctx, &ctx->cell_of, (size_t)P * ctx->n_mols));
        CHK(dalloc(ctx, &ctx->crank, (size_t)P * ctx->n_mols));
        CHK(dalloc(ctx, &ctx->member, (size_t)P * ctx->n_mols));
        CHK(dalloc(ctx, &ctx->cell_pos, (size_t)P * ctx->n_mols));
        CHK(dalloc(ctx, &ctx->cstart, (size_t)P * (ctx->cg.ncell + 1)));
        const size_t nbm = (size_t)P * ctx->n_mols * ((ctx->n_mols + 63) / 64);
        if (nbm * sizeof(unsigned long long) <= NL_BM_MAX_BYTES) CHK(dalloc(ctx, &ctx->nl_bm, nbm));
    }
    CHK(dalloc(ctx, &ctx->flag, P));
    CHK(dalloc(ctx, &ctx->status, 1));
    CHK(dalloc(ctx, &ctx->nreb, P));
    CHK(dalloc(ctx, &ctx->qpm, (size_t)E * N3));
    CHK(dalloc(ctx, &ctx->vpart, (size_t)E * 6 * ctx->ep_cap));
    CHK(dalloc(ctx, &ctx->vout, (size_t)E * 6));
    CHK(dalloc(ctx, &ctx->rpart, (size_t)E * ctx->nblk * N3));
    if (ctx->rs_cull) {
        CHK(dalloc(ctx, &ctx->rs_sites, (size_t)E * ctx->nblk * 64));
        CHK(dalloc(ctx, &ctx->rs_box, (size_t)E * ctx->nblk * RS_BOX));
        CHK(dalloc(ctx, &ctx->rs_far, (size_t)E * ctx->ntile));
    }
    CHK(dalloc(ctx, &ctx->spart, (size_t)E * ctx->nac * ctx->G * 8));
    CHK(dalloc(ctx, &ctx->sprime, (size_t)E * ctx->G * KF_REC));

    CHK(dalloc(ctx, &ctx->flj, (size_t)E * N3));
    CHK(dalloc(ctx, &ctx->fspl, (size_t)E * N3));
    CHK(dalloc(ctx, &ctx->fpm, (size_t)E * N3));
    if (ctx->kf_split_ok) CHK(dalloc(ctx, &ctx->fk2, (size_t)E * N3));
    if (ctx->kf_mf_ok) CHK(dalloc(ctx, &ctx->wm, (size_t)E * 4 * ctx->mf_nks * ctx->mf_Mp));
    CHK(dalloc(ctx, &ctx->Ve, E));
    CHK(dalloc(ctx, &ctx->epart, (size_t)E * ctx->ep_cap));
    CHK(dalloc(ctx, &ctx->scal, (size_t)E * 4));
    CHK(dalloc(ctx, &ctx->out, 8));
    CHK(dalloc(ctx, &ctx->heat, 1));
    CHK(dalloc(ctx, &ctx->heat_step, 1));
    CHK(dalloc(ctx, &ctx->step_ctr, 1));
    CHK(dalloc(ctx, &ctx->cmf_acc, (size_t)N3));
    CHK(dalloc(ctx, &ctx->cmf_n, 1));
    CHK(dalloc(ctx, &ctx->qc, (size_t)N3));
    CHK(dalloc(ctx, &ctx->nq, (size_t)P * N3));
    CHK(dalloc(ctx, &ctx->np, (size_t)P * N3));
    CHK(dalloc(ctx, &ctx->ppart, 2 * (size_t)ctx->ptile + 2));
    CHK(dalloc(ctx, &ctx->hpart, (size_t)ctx->P * ctx->ptile + 1));
    CHK(dalloc(ctx, &ctx->obs, (size_t)ctx->ptile * 4 + 64));
    CHK(dalloc(ctx, &ctx->obs_out, 4));
    CHK(dalloc(ctx, &ctx->ntrace, 1));
    ctx->dev_cap = 3 + (long)(P > 1 ? P - 1 : 0) * ctx->n_dof;
    CHK(dalloc(ctx, &ctx->dev, (size_t)ctx->dev_cap));
    return QTP_OK;
}

// Per-mode propagator and thermostat constants {scale, w^2, c1, c2 * sqrt(mass
// factor)} and propagation-mass factors {f, sqrt f} (update.cpp:302-304,
// thermostat.cpp:119-121).  Partially adiabatic CMD (pimd_mode 1) moves every
// non-centroid mode to the frequency W with mass m (w_k / W)^2 (the spring
// m w_k^2 is unchanged) and damps it critically, c1 = exp(-2 W dt).
static int upload_modes(qtp_ctx *ctx)
{
    const int P = ctx->P;
    std::vector<double4> md(P);
    std::vector<double2> ms(P, make_double2(1.0, 1.0));
    const double pre = 2.0 / (1.0 * ctx->drv.beta_n), pre_sq = pre * pre;
    const double tpre = 4.0 / (ctx->drv.beta_n * 1.0);
    const double dt = ctx->drv.dt;
    const double W = ctx->omega_ad;
    for (int k = 0; k < P; ++k) {
        const double s = std::sin(k * ctx->drv.piovern);              // update.cpp:302
        const double wsq = pre_sq * s * s;
        if (ctx->pimd_mode == 1 && k > 0) {
            const double f = wsq / (W * W);
            const double c1 = std::exp(-2.0 * W * dt);
            md[k] = make_double4(1.0 / std::sqrt(1.0 + 0.25 * W * W * dt * dt), W * W, c1,
                                 std::sqrt(1.0 - c1 * c1) * std::sqrt(f));
            ms[k] = make_double2(f, std::sqrt(f));
            continue;
        }
        const double scale = 1.0 / std::sqrt(1.0 + 0.25 * wsq * dt * dt);
        const double gk = tpre * std::sin(k * RC.pi / P);              // thermostat.cpp:119
        const double c1 = std::exp(-dt * gk);
        const double c2 = std::sqrt(1.0 - c1 * c1);
        md[k] = make_double4(scale, wsq, c1, c2);
    }
    HIPCHK(hipMemcpy(ctx->mode, md.data(), P * sizeof(double4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->ms, ms.data(), P * sizeof(double2), hipMemcpyHostToDevice));
    return QTP_OK;
}

static int upload_constants(qtp_ctx *ctx)
{
    // masses and charges (simulation.cpp:811-822)
    ctx->h_mass_atom.assign(ctx->NAP, 1.0);
    ctx->h_charge.assign(ctx->NAP, 0.0);
    std::vector<double> sqm(ctx->NAP, 0.0);
    for (int i = 0; i < ctx->n_mols; ++i) {
        ctx->h_mass_atom[3 * i] = RC.mo;
        ctx->h_mass_atom[3 * i + 1] = RC.mh;
        ctx->h_mass_atom[3 * i + 2] = RC.mh;
        ctx->h_charge[3 * i] = RC.qo;
        ctx->h_charge[3 * i + 1] = RC.qh;
        ctx->h_charge[3 * i + 2] = RC.qh;
    }
    for (int a = 0; a < ctx->n_atoms; ++a) sqm[a] = std::sqrt(ctx->h_mass_atom[a] / ctx->drv.beta_n);
    HIPCHK(hipMemcpy(ctx->m_atom, ctx->h_mass_atom.data(), ctx->NAP * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->sqm_atom, sqm.data(), ctx->NAP * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->charge, ctx->h_charge.data(), ctx->NAP * sizeof(double), hipMemcpyHostToDevice));
    ctx->eself = self_energy(ctx, ctx->h_charge.data());
    // normal modes and per-mode propagator / thermostat constants
    const int P = ctx->P;
    std::vector<double> C;
    build_nm(P, C);
    HIPCHK(hipMemcpy(ctx->Cm, C.data(), C.size() * sizeof(double), hipMemcpyHostToDevice));
    return upload_modes(ctx);
}


