// qtp_host_run.hpp -- host side of the gfx950 qTIP4P/F step: launch
// sequences (force phases, SC forces, neighbour lists, the SC and classical
// steps) and host <-> device layout conversion.
// Included once, in order, by qtip4pf.hip (one translation unit).
#pragma once

// ---------------------------------------------------------------------------
// force phase: E single-bead evaluations (forces.cpp:468-630)
// ---------------------------------------------------------------------------
// Launch shape of k_ksum_mf for E evaluations: blocks per chunk group (a
// multiple of the count that keeps KS_MAXU units per wave; more when the
// launch has few blocks, or as tuned), chunks per block (as tuned, else the
// most of 1, 2, 4, 8 that still leaves 2048 blocks), chunk groups.
static void ksum_mf_shape(const qtp_ctx *ctx, int E, int *usplit, int *cpb, int *ncg)
{
    const int per = KS_MAXU * KS_WAVES;
    const int req = std::max(1, (ctx->ks_nunit + per - 1) / per);
    const int umax = std::min(4, std::max(1, ctx->ks_nunit / (req * KS_WAVES)));
    const int mult = ctx->ks_usplit > 0 ? std::min(ctx->ks_usplit, umax)
                                        : ((long)E * ctx->nac * req < 128 ? umax : 1);
    *usplit = req * mult;
    int c = ctx->ks_cpb;
    if (c <= 0) {
        c = 1;
        for (int t = 2; t <= 8; t *= 2)
            if ((long)E * ((ctx->nac + t - 1) / t) * *usplit >= 2048) c = t;
    }
    *cpb = c;
    *ncg = (ctx->nac + c - 1) / c;
}

static int run_lists(qtp_ctx *ctx, int force_all, bool with_heat = false,
                     const unsigned long long *ctr = nullptr, hipStream_t st = nullptr);

// energy partials summed per evaluation: the spline blocks only when used
static int ep_live(const qtp_ctx *ctx) { return ctx->spl_on ? ctx->ep_stride : ctx->ep_spl; }

// The splined potential is part of a force call only through ctx->spl_on,
// which every entry point sets for its duration.
struct SplineScope {
    qtp_ctx *ctx;
    bool prev;
    SplineScope(qtp_ctx *c, bool on) : ctx(c), prev(c->spl_on) { c->spl_on = on; }
    ~SplineScope() { ctx->spl_on = prev; }
};


// vout (optional): per-evaluation virial W_ab = sum_i r_i (x) F_i in the
// minimum-image pair form (6 components xx yy zz xy xz yz), at fixed Ewald
// parameters.
// fout2: when given, k-space forces are split in two halves (k_kforce's split
// form) and the second half goes there; the caller adds it.
static int run_phase(qtp_ctx *ctx, const double *pos, int E, int lmode, int lbase, double *fout,
                     double *eout, double *vout = nullptr, double *fout2 = nullptr)
{
    hipStream_t s = ctx->stream;
    const long N3 = 3L * ctx->NAP;
    const double g = ctx->ff.gamma_vs, h = ctx->ff.half1mg;
    double *vp = vout ? ctx->vpart : nullptr;
    // The k-space sums (ksum, sreduce) and the LJ/intramolecular kernel do not
    // depend on the real-space kernel: they run on the second stream, joined
    // before kforce, with a share of the real-space tiles after them.  With
    // per-kernel timing on, everything stays on one stream so each launch is
    // timed alone.
    const bool fork = !ctx->timing;
    hipStream_t sk = fork ? ctx->s2 : s;
    // real-space tiles: the first nt_a on this stream, the other nt_b (a
    // tunable share, in eighths) on the second stream after the LJ kernel,
    // where its kernels leave the chip room beside the first part
    const int nt_b = fork ? (int)((long)ctx->ntile * ctx->rs_split / 8) : 0;
    const int nt_a = ctx->ntile - nt_b;
    // coefficients in SGPRs (more waves per SIMD) for large grids, where
    // that measured faster (4096 x 32: 3.42 against 3.60 ms); literals
    // for small ones (216 x 32: no gain, and the second stream's kernels
    // lost more than the launch gained)
    const bool carg = ctx->rs_carg >= 0 ? ctx->rs_carg == 1
                                        : (long)E * ctx->ntile >= RS_CARG_BLOCKS;
    const RsCoef cf = rs_coef();
    auto launch_rs = [&](hipStream_t st, int t0, int nt) {
        TimeScope ts(ctx, "realspace");
#define QTP_RS_LAUNCH_W(VIR, CARG, CULL, W)                                                        \
    k_realspace<VIR, CARG, CULL, W><<<E * nt, 64 * W, 0, st>>>(                                    \
        pos, N3, ctx->charge, ctx->n_atoms, ctx->NAP, ctx->nblk, g, h, ctx->tiles, ctx->ntile,     \
        t0, nt, ctx->box, ctx->ew, ctx->rpart, ctx->epart, ctx->ep_stride, ctx->ep_real, vp, cf, \
        ctx->rs_sites, ctx->rs_box, ctx->rs_far)
#define QTP_RS_LAUNCH(VIR, CARG, CULL)                                                             \
    do {                                                                                           \
        if (ctx->rs_waves == 2) QTP_RS_LAUNCH_W(VIR, CARG, CULL, 2);                              \
        else QTP_RS_LAUNCH_W(VIR, CARG, CULL, 4);                                                  \
    } while (0)
        if (ctx->rs_cull) {
            if (vout) QTP_RS_LAUNCH(true, true, true);
            else QTP_RS_LAUNCH(false, true, true);
        } else if (vout) {
            if (carg) QTP_RS_LAUNCH(true, true, false);
            else QTP_RS_LAUNCH(true, false, false);
        } else {
            if (carg) QTP_RS_LAUNCH(false, true, false);
            else QTP_RS_LAUNCH(false, false, false);
        }
#undef QTP_RS_LAUNCH
#undef QTP_RS_LAUNCH_W
    };
    // the culling form's site tables and skip flags, ahead of the fork: on
    // their own the two small kernels take a few microseconds, behind the
    // structure-factor blocks of the other stream many times that
    if (ctx->rs_cull) {
        {
            TimeScope ts(ctx, "sites");
            k_sites<<<E * ctx->nblk, 64, 0, s>>>(pos, N3, ctx->charge, ctx->n_atoms, ctx->NAP,
                                                 ctx->nblk, g, h, ctx->rs_sites, ctx->rs_box);
        }
        TimeScope ts(ctx, "tilefar");
        const long nk = (long)E * ctx->ntile;
        k_tilefar<<<(unsigned)((nk + 255) / 256), 256, 0, s>>>(
            ctx->tiles, ctx->ntile, ctx->nblk, E, ctx->rs_box, ctx->box, ctx->ew, ctx->rs_far,
            ctx->epart, ctx->ep_stride, ctx->ep_real, vp);
    }
    if (fork) {
        HIPCHK(hipEventRecord(ctx->ev_fork, s));
        HIPCHK(hipStreamWaitEvent(sk, ctx->ev_fork, 0));
    }
    {
        TimeScope ts(ctx, "ksum");
        if (ctx->kf_mf) {
            int usplit, cpb, ncg;
            ksum_mf_shape(ctx, E, &usplit, &cpb, &ncg);
            k_ksum_mf<<<E * ncg * usplit, 64 * KS_WAVES, ctx->ksum_mf_lds, sk>>>(
                pos, N3, ctx->charge, ctx->n_atoms, ctx->NAP, g, h, ctx->ew, ctx->ks_col,
                ctx->ks_ncol4, ctx->ks_unit, ctx->ks_nunit, ctx->G, ctx->AC, ctx->nac, cpb, usplit,
                ctx->spart);
        } else
            k_ksum<<<E * ctx->ngc * ctx->nac, 256, ctx->ksum_lds, sk>>>(
                pos, N3, ctx->charge, ctx->n_atoms, ctx->NAP, g, h, ctx->ew, ctx->groups, ctx->G,
                ctx->AC, ctx->ngc, ctx->nac, ctx->spart);
    }
    {
        TimeScope ts(ctx, "sreduce");
        // partials per group: one per chunk (vector k_ksum) or chunk group
        int nparts = ctx->nac;
        if (ctx->kf_mf) {
            int usplit, cpb;
            ksum_mf_shape(ctx, E, &usplit, &cpb, &nparts);
        }
        k_sreduce<<<E * ctx->ngr, 256, 0, sk>>>(ctx->spart, nparts, ctx->G, ctx->groups, ctx->gcf,
                                                 ctx->sprime, ctx->epart, ctx->ep_stride, ctx->ep_k,
                                                 ctx->grk, 0.25 / ctx->ew.kappa2, vp,
                                                 ctx->kf_mf ? ctx->wm : nullptr,
                                                 ctx->mf_Mp, 4 * ctx->mf_nks, ctx->gcol,
                                                 ctx->ew.bk[2], ctx->kf_mf ? 1 : 0);
    }
    // a list check deferred by the step (only the LJ kernel reads the lists)
    if (ctx->lists_defer) {
        ctx->lists_defer = false;
        CHK(run_lists(ctx, 0, ctx->lists_heat, ctx->lists_ctr, sk));
    }
    {
        TimeScope ts(ctx, "ljintra");
#define QTP_LJ_LAUNCH(VIR, LPM)                                                                    \
    k_ljintra<VIR, LPM><<<E * ctx->nmb, 256, 0, sk>>>(pos, N3, ctx->n_mols, ctx->NAP, ctx->fpoint, \
                                                      ctx->flist, ctx->fcap, lmode, lbase,         \
                                                      ctx->box, ctx->ff, ctx->flj, ctx->epart,     \
                                                      ctx->ep_stride, ctx->ep_lj, vp)
        if (ctx->lj_lpm == 64) {
            if (vout) QTP_LJ_LAUNCH(true, 64);
            else QTP_LJ_LAUNCH(false, 64);
        } else {
            if (vout) QTP_LJ_LAUNCH(true, 16);
            else QTP_LJ_LAUNCH(false, 16);
        }
#undef QTP_LJ_LAUNCH
    }
    if (ctx->spl_on) {
        TimeScope ts(ctx, "spline");
        if (vout)
            k_spline<true><<<E * ctx->nblk, 256, 0, sk>>>(
                pos, N3, ctx->n_atoms, ctx->NAP, ctx->nblk, ctx->box, ctx->spl, ctx->spl_coef,
                ctx->fspl, ctx->epart, ctx->ep_stride, ctx->ep_spl, vp);
        else
            k_spline<false><<<E * ctx->nblk, 256, 0, sk>>>(
                pos, N3, ctx->n_atoms, ctx->NAP, ctx->nblk, ctx->box, ctx->spl, ctx->spl_coef,
                ctx->fspl, ctx->epart, ctx->ep_stride, ctx->ep_spl, vp);
    }
    if (nt_b > 0) launch_rs(sk, nt_a, nt_b);
    if (fork) HIPCHK(hipEventRecord(ctx->ev_join, sk));
    launch_rs(s, 0, nt_a);
    if (fork) HIPCHK(hipStreamWaitEvent(s, ctx->ev_join, 0));
    {
        TimeScope ts(ctx, "kforce");
        // tiles the CULL form skipped: their (zero) slots are not read
        const unsigned char *rfar = ctx->rs_cull ? ctx->rs_far : nullptr;
        const int nab = (ctx->n_atoms + KF_SITES - 1) / KF_SITES;
        const int nbm = E * ((ctx->n_atoms + MF_SITES - 1) / MF_SITES);
        // few blocks: twice the waves per block (when its LDS fits the default 64 KB)
        const bool small = ctx->mf_small >= 0 ? ctx->mf_small == 1 : nbm < MF_SMALL_BLOCKS;
        const bool hi = ctx->mf_nks > MF_KS_LO;
#define QTP_KF_MF(W, KS, LDS, RT, NRT)                                                         \
    k_kforce_mf<W, KS><<<nbm, 64 * W, LDS, s>>>(pos, N3, ctx->charge, ctx->n_atoms, ctx->NAP, g, h, \
                                                ctx->ew, ctx->wm, ctx->mf_Mp, ctx->mf_nks, RT, NRT, \
                                                ctx->rpart, ctx->nblk, ctx->flj,                  \
                                                ctx->spl_on ? ctx->fspl : nullptr, fout, rfar,    \
                                                ctx->ntile)
        if (ctx->kf_mf && small && ctx->kforce_mf_lds_small <= 64 * 1024) {
            if (hi) QTP_KF_MF(MF_WAVES_SMALL, MF_MAXKS, ctx->kforce_mf_lds_small, ctx->mf_rt_small, ctx->mf_nrt_small);
            else QTP_KF_MF(MF_WAVES_SMALL, MF_KS_LO, ctx->kforce_mf_lds_small, ctx->mf_rt_small, ctx->mf_nrt_small);
        } else if (ctx->kf_mf) {
            if (hi) QTP_KF_MF(MF_WAVES, MF_MAXKS, ctx->kforce_mf_lds, ctx->mf_rt, ctx->mf_nrt);
            else QTP_KF_MF(MF_WAVES, MF_KS_LO, ctx->kforce_mf_lds, ctx->mf_rt, ctx->mf_nrt);
        }
#undef QTP_KF_MF
        if (!ctx->kf_mf) {
            if (fout2)
                k_kforce<KF_SPLIT_WAVES, true><<<2 * E * nab, 64 * KF_SPLIT_WAVES, ctx->kforce_lds_split, s>>>(
                    pos, N3, ctx->charge, ctx->n_atoms, ctx->NAP, g, h, ctx->ew, ctx->G, ctx->sprime,
                    ctx->rpart, ctx->nblk, ctx->flj, ctx->spl_on ? ctx->fspl : nullptr, fout, fout2,
                    ctx->kf_chunk_split, rfar, ctx->ntile);
            else
                k_kforce<KF_WAVES, false><<<E * nab, 64 * KF_WAVES, ctx->kforce_lds, s>>>(
                    pos, N3, ctx->charge, ctx->n_atoms, ctx->NAP, g, h, ctx->ew, ctx->G, ctx->sprime,
                    ctx->rpart, ctx->nblk, ctx->flj, ctx->spl_on ? ctx->fspl : nullptr, fout, nullptr,
                    ctx->kf_chunk, rfar, ctx->ntile);
        }
    }
    if (eout) {
        TimeScope ts(ctx, "esum");
        k_esum<<<E, 64, 0, s>>>(E, ctx->epart, ctx->ep_stride, ep_live(ctx), ctx->eself, eout);
    }
    if (vout) k_vsum<<<E, 64, 0, s>>>(ctx->vpart, ctx->ep_stride, ep_live(ctx), vout);
    HIPCHK(hipGetLastError());
    return QTP_OK;
}

static SCArgs sc_args(qtp_ctx *ctx)
{
    SCArgs a;
    a.P = ctx->P;
    a.n_atoms = ctx->n_atoms;
    a.NAP = ctx->NAP;
    a.omega_psq = ctx->drv.omega_psq;
    a.epsilonsc = ctx->drv.epsilonsc;
    a.halfdt = ctx->drv.halfdt;
    return a;
}

// SC forces for all beads from ctx->q with the current lists (forces.cpp:698-789)
static int run_sc_forces(qtp_ctx *ctx, int kick, unsigned long long *ctr)
{
    hipStream_t s = ctx->stream;
    const int P = ctx->P;
    double *f2 = ctx->kf_split ? ctx->fk2 : nullptr;
    CHK(run_phase(ctx, ctx->q, P, 0, 0, ctx->fph, nullptr, nullptr, f2));
    SCArgs a = sc_args(ctx);
    {
        TimeScope ts(ctx, "sc1");
        k_sc1<<<P, SC_THREADS, 0, s>>>(a, ctx->q, ctx->fph, f2, ctx->epart, ctx->ep_stride, ep_live(ctx), ctx->eself,
                                ctx->Ve, ctx->m_atom, ctx->F, ctx->qpm, ctx->scal);
    }
    CHK(run_phase(ctx, ctx->qpm, P, 1, 0, ctx->fpm, nullptr, nullptr, f2));
    {
        TimeScope ts(ctx, "sc2");
        // at most four runs of SC_THREADS dofs per block (one block per bead
        // up to 455 molecules)
        const int nch = std::max(1, (3 * ctx->NAP + 4 * SC_THREADS - 1) / (4 * SC_THREADS));
        k_sc2<<<P * nch + 1, SC_THREADS, 0, s>>>(a, ctx->fph, ctx->fpm, f2, ctx->scal, ctx->F, ctx->p, kick, ctx->out,
                                ctr, nch);
    }
    HIPCHK(hipGetLastError());
    return QTP_OK;
}

// with_heat: also reduce a thermostatted step's heat (folded into the fused
// list kernel; a kernel of its own on the multi-kernel path)
// Cell-binned build of the beads flagged in ctx->flag (check = 0) or of
// those the displacement test flags (check = 1): bins, row counts, offsets,
// rows.
static int run_cell_lists(qtp_ctx *ctx, int check, int force_all,
                          const unsigned long long *ctr)
{
    hipStream_t s = ctx->stream;
    const int P = ctx->P, nrb = (ctx->n_mols + 3) / 4;
    const size_t bm_lds = 4 * (size_t)((ctx->n_mols + 63) / 64) * sizeof(unsigned long long);
    {
        TimeScope ts(ctx, "check");
        k_cell_build<<<P, 1024, ctx->cg.ncell * sizeof(int), s>>>(
            ctx->n_mols, ctx->NAP, ctx->NMP, ctx->q, ctx->box, ctx->cg,
            ctx->drv.r_skin - ctx->drv.lj_rcut, check, force_all, ctr ? ctx->rebuild_every : 0, ctr,
            ctx->disp, ctx->flag, check ? ctx->nreb : nullptr, ctx->cell_of, ctx->crank,
            ctx->cstart, ctx->member, ctx->cell_pos);
    }
    {
        TimeScope ts(ctx, "nlist");
        k_nl_cell<false><<<P * nrb, 256, bm_lds, s>>>(
            ctx->n_mols, ctx->NAP, ctx->q, ctx->box, ctx->cg, ctx->ff.r_skin_sq, ctx->list_size,
            ctx->fcap, ctx->flag, ctx->cell_of, ctx->cstart, ctx->member, ctx->cell_pos, ctx->cnt,
            nullptr, nullptr, nullptr, nullptr, ctx->nl_bm);
        k_nl_scan<<<P, 1024, 0, s>>>(ctx->n_mols, ctx->list_size, ctx->fcap, ctx->flag, ctx->cnt,
                                     ctx->point, ctx->fpoint, ctx->list, ctx->disp, ctx->NMP,
                                     ctx->status);
        k_nl_cell<true><<<P * nrb, 256, bm_lds, s>>>(
            ctx->n_mols, ctx->NAP, ctx->q, ctx->box, ctx->cg, ctx->ff.r_skin_sq, ctx->list_size,
            ctx->fcap, ctx->flag, ctx->cell_of, ctx->cstart, ctx->member, ctx->cell_pos, nullptr,
            ctx->point, ctx->fpoint, ctx->list, ctx->flist, ctx->nl_bm);
    }
    HIPCHK(hipGetLastError());
    return QTP_OK;
}

// ctr: the device step counter of resident steps (forced rebuilds every
// ctx->rebuild_every steps), or null
static int run_lists(qtp_ctx *ctx, int force_all, bool with_heat, const unsigned long long *ctr,
                     hipStream_t st)
{
    const int every = ctr ? ctx->rebuild_every : 0;
    hipStream_t s = st ? st : ctx->stream;
    const int P = ctx->P;
    HeatArgs heat{};
    if (with_heat) {
        heat.part = ctx->ppart;
        heat.hpart = ctx->hpart;
        heat.heat_acc = ctx->heat;
        heat.heat_step = ctx->heat_step;
        heat.ntile = ctx->ptile;
        heat.nh = ctx->P * ctx->ptile;
    }
    const bool cells = use_cells(ctx);
    if (ctx->n_mols <= NL_FUSED_MAXMOL && !cells) {
        TimeScope ts(ctx, "nlist");
        k_nlist_bead<<<P + (with_heat ? 1 : 0), 1024, 0, s>>>(ctx->n_mols, ctx->NAP, ctx->NMP, ctx->q, ctx->box,
                                        ctx->ff.r_skin_sq, ctx->drv.r_skin - ctx->drv.lj_rcut, force_all,
                                        every, ctr, ctx->list_size, ctx->fcap, ctx->disp, ctx->flag, ctx->nreb,
                                        ctx->point, ctx->fpoint, ctx->list, ctx->flist,
                                        ctx->status, heat);
        HIPCHK(hipGetLastError());
        return QTP_OK;
    }
    if (with_heat)
        k_heat<<<1, 1024, 0, s>>>(heat.ntile, heat.nh, heat.part, heat.hpart, heat.heat_acc,
                                heat.heat_step);
    if (cells) return run_cell_lists(ctx, 1, force_all, ctr);
    {
        TimeScope ts(ctx, "check");
        k_check<<<P, 256, 0, s>>>(ctx->n_mols, ctx->NMP, ctx->disp, ctx->drv.r_skin - ctx->drv.lj_rcut, force_all,
                                  every, ctr, ctx->flag, ctx->nreb);
    }
    const int nrb = (ctx->n_mols + 3) / 4;
    {
        TimeScope ts(ctx, "nlist");
        k_nl_count<<<P * nrb, 256, 0, s>>>(ctx->n_mols, ctx->NAP, ctx->q, ctx->box, ctx->ff.r_skin_sq,
                                           ctx->flag, ctx->cnt);
        k_nl_scan<<<P, 1024, 0, s>>>(ctx->n_mols, ctx->list_size, ctx->fcap, ctx->flag, ctx->cnt,
                                     ctx->point, ctx->fpoint, ctx->list, ctx->disp, ctx->NMP,
                                     ctx->status);
        k_nl_fill<<<P * nrb, 256, 0, s>>>(ctx->n_mols, ctx->NAP, ctx->q, ctx->box, ctx->ff.r_skin_sq,
                                          ctx->list_size, ctx->fcap, ctx->flag, ctx->point,
                                          ctx->fpoint, ctx->list, ctx->flist);
    }
    HIPCHK(hipGetLastError());
    return QTP_OK;
}

static PropArgs prop_args(qtp_ctx *ctx, int nvt, int host_dev)
{
    PropArgs A;
    A.P = ctx->P;
    A.n_atoms = ctx->n_atoms;
    A.NAP = ctx->NAP;
    A.ntile = ctx->ptile;
    A.halfdt = ctx->drv.halfdt;
    A.dt = ctx->drv.dt;
    A.beta_n = ctx->drv.beta_n;
    A.nvt = nvt;
    A.host_dev = host_dev;
    A.n_dof = ctx->n_dof;
    A.seed = ctx->seed;
    A.cmode = ctx->pimd_mode == 2 ? 2 : 0;
    A.cthermo = ctx->cthermo;
    double c = ctx->tau_fs > 0.0 ? std::exp(-ctx->prm.dtfs / ctx->tau_fs) : 0.0;
    if (host_dev) {   // drop-in steps: the reference's pile_g (tau -> 0)
        A.cthermo = 0;
        c = 0.0;
    }
    A.cbussi = A.cthermo == 0 ? c : 0.0;
    A.c1_0 = c;
    A.c2_0 = std::sqrt(1.0 - c * c);
    return A;
}

// One SC step on the device state (update.cpp:271-401).
static int run_sc_step(qtp_ctx *ctx, int nvt, int host_dev, unsigned long long *ctr)
{
    hipStream_t s = ctx->stream;
    PropArgs A = prop_args(ctx, nvt, host_dev);
    const int grid = ctx->ptile * ctx->P;
    // one kernel per tile, the beads staged in LDS, when the centroid needs
    // no rescale over the whole system between the two transforms
    const bool bussi = nvt && A.cthermo == 0 && A.cmode != 2;
    const bool big = ctx->ptile >= PT_MIN_TILES;
    if ((ctx->prop_fused == 1 || (ctx->prop_fused == -1 && big)) && !bussi &&
        !(nvt && host_dev) && ctx->P <= PT_WAVES * PT_PER_WAVE) {
        TimeScope ts(ctx, "prop");
        k_prop_tile<<<ctx->ptile, 64 * PT_WAVES, prop_tile_lds(ctx->P), s>>>(
            A, ctx->Cm, ctx->mode, ctx->ms, ctx->q, ctx->p, ctx->F, ctx->m_atom, ctx->sqm_atom,
            ctx->ppart, ctx->hpart, ctx->disp, ctx->NMP, ctr);
    } else {
    {
        TimeScope ts(ctx, "prop_a");
        k_prop_fwd<<<grid, 64, 0, s>>>(A, ctx->Cm, ctx->mode, ctx->ms, ctx->q, ctx->p, ctx->F, ctx->m_atom,
                                       ctx->sqm_atom, ctx->dev, ctx->nq, ctx->np, ctx->ppart,
                                       ctx->hpart, ctr);
    }
    {
        TimeScope ts(ctx, "prop_b");
        k_prop_inv<<<grid, 64, 0, s>>>(A, ctx->Cm, ctx->mode, ctx->nq, ctx->np, ctx->m_atom,
                                       ctx->ppart, ctx->dev, ctx->q, ctx->p, ctx->disp, ctx->NMP,
                                       ctx->hpart, ctr);
    }
    }
    HIPCHK(hipGetLastError());
    // the one-kernel list check and rebuild: on the first force phase's
    // second stream, ahead of the LJ kernel, so that the real-space kernel
    // starts right after the propagation
    if (ctx->n_mols <= NL_FUSED_MAXMOL && !use_cells(ctx)) {
        ctx->lists_defer = true;
        ctx->lists_heat = nvt != 0;
        ctx->lists_ctr = ctr;
    } else {
        CHK(run_lists(ctx, 0, nvt != 0, ctr));
    }
    const int st = run_sc_forces(ctx, 1, ctr);
    ctx->lists_defer = false;  // consumed by the first phase (also on an error)
    CHK(st);
    if (ctx->cmf_on)
        k_cmf_acc<<<(3 * ctx->NAP + 255) / 256, 256, 0, s>>>(ctx->P, ctx->NAP, ctx->F, ctx->cmf_acc,
                                                            ctx->cmf_n);
    HIPCHK(hipGetLastError());
    return QTP_OK;
}

// One classical velocity-Verlet step (update.cpp:78-137).  Resident steps
// (resident) add the learned correction when a model is attached and, when
// nvt, a Langevin step on every atom first.
static int run_class_step(qtp_ctx *ctx, unsigned long long *ctr, bool resident = false,
                          int nvt = 0)
{
    hipStream_t s = ctx->stream;
    const int nb = (3 * ctx->NAP + 255) / 256;
    VVLang lang{};
    if (resident && nvt) {
        const double c = std::exp(-ctx->prm.dtfs / ctx->tau_fs);
        lang.on = 1;
        lang.c1 = c;
        lang.c2 = std::sqrt(1.0 - c * c);
        lang.seed = ctx->seed;
    }
    {
        TimeScope ts(ctx, "vv_a");
        k_vv_a<<<nb, 256, 0, s>>>(ctx->n_atoms, ctx->NAP, ctx->NMP, ctx->drv.halfdt, ctx->drv.dt,
                                  ctx->m_atom, ctx->F, ctx->q, ctx->p, ctx->disp, lang,
                                  ctx->sqm_atom, ctx->step_ctr);
    }
    // the learned correction depends on the new positions only: it runs on a
    // third stream beside the lists and the qTIP4P/F forces (on the main
    // stream when kernels are timed one by one)
    const bool with_nn = resident && ctx->nn;
    const bool nn_fork = with_nn && !ctx->timing;
    if (nn_fork) {
        HIPCHK(hipEventRecord(ctx->ev_nn_fork, s));
        HIPCHK(hipStreamWaitEvent(ctx->s3, ctx->ev_nn_fork, 0));
        CHK(run_nn_forces(ctx, ctx->s3));
        HIPCHK(hipEventRecord(ctx->ev_nn_join, ctx->s3));
    }
    CHK(run_lists(ctx, 0, false, resident ? ctr : nullptr));
    CHK(run_phase(ctx, ctx->q, 1, 0, 0, ctx->fph, nullptr));
    if (with_nn && !nn_fork) CHK(run_nn_forces(ctx, s));
    if (nn_fork) HIPCHK(hipStreamWaitEvent(s, ctx->ev_nn_join, 0));
    {
        TimeScope ts(ctx, "vv_b");
        k_vv_b<<<nb, 256, 0, s>>>(ctx->n_atoms, ctx->NAP, ctx->drv.halfdt, ctx->fph, ctx->epart,
                                  ctx->ep_stride, ep_live(ctx), ctx->eself, ctx->F, ctx->p, ctx->out,
                                  ctr, with_nn ? ctx->fnn : nullptr,
                                  with_nn ? ctx->nn->Etile : nullptr, with_nn ? nn_tiles(ctx->nn) : 0);
    }
    HIPCHK(hipGetLastError());
    return QTP_OK;
}

// Forces (with the learned correction when attached) and energy of the
// resident classical state into F and out[0], no step: the set-up of
// qtp_set_state / qtp_load_checkpoint.
static int class_initial_forces(qtp_ctx *ctx)
{
    CHK(run_phase(ctx, ctx->q, 1, 0, 0, ctx->fph, nullptr));
    if (ctx->nn) CHK(run_nn_forces(ctx, ctx->stream));
    const int nb = (3 * ctx->NAP + 255) / 256;
    k_vv_b<<<nb, 256, 0, ctx->stream>>>(ctx->n_atoms, ctx->NAP, 0.0, ctx->fph, ctx->epart,
                                        ctx->ep_stride, ep_live(ctx), ctx->eself, ctx->F, ctx->p,
                                        ctx->out, nullptr, ctx->nn ? ctx->fnn : nullptr,
                                        ctx->nn ? ctx->nn->Etile : nullptr,
                                        ctx->nn ? nn_tiles(ctx->nn) : 0);
    HIPCHK(hipGetLastError());
    return QTP_OK;
}

// ---------------------------------------------------------------------------
// host <-> device layout conversion
// ---------------------------------------------------------------------------
static void to_planar(const qtp_ctx *ctx, const double *ref, int nb, std::vector<double> &out)
{
    const long N3 = 3L * ctx->NAP;
    out.assign((size_t)nb * N3, 0.0);
    for (int b = 0; b < nb; ++b)
        for (int a = 0; a < ctx->n_atoms; ++a)
            for (int d = 0; d < 3; ++d)
                out[b * N3 + (long)d * ctx->NAP + a] = ref[(long)b * ctx->n_dof + 3 * a + d];
}

static void from_planar(const qtp_ctx *ctx, const std::vector<double> &pl, int nb, double *ref)
{
    const long N3 = 3L * ctx->NAP;
    for (int b = 0; b < nb; ++b)
        for (int a = 0; a < ctx->n_atoms; ++a)
            for (int d = 0; d < 3; ++d)
                ref[(long)b * ctx->n_dof + 3 * a + d] = pl[b * N3 + (long)d * ctx->NAP + a];
}

static int upload_planar(qtp_ctx *ctx, double *dst, const double *ref, int nb)
{
    std::vector<double> h;
    to_planar(ctx, ref, nb, h);
    HIPCHK(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return QTP_OK;
}

static int download_planar(qtp_ctx *ctx, const double *src, int nb, double *ref)
{
    std::vector<double> h((size_t)nb * 3 * ctx->NAP);
    HIPCHK(hipMemcpyAsync(h.data(), src, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    from_planar(ctx, h, nb, ref);
    return QTP_OK;
}

// disp: reference [b][3*mol + d]  <->  device [b][d][NMP]
static int upload_disp(qtp_ctx *ctx, const double *ref, int b0, int nb)
{
    std::vector<double> h((size_t)nb * 3 * ctx->NMP, 0.0);
    for (int b = 0; b < nb; ++b)
        for (int i = 0; i < ctx->n_mols; ++i)
            for (int d = 0; d < 3; ++d)
                h[((size_t)b * 3 + d) * ctx->NMP + i] = ref[(long)(b0 + b) * 3 * ctx->n_mols + 3 * i + d];
    HIPCHK(hipMemcpy(ctx->disp + (size_t)b0 * 3 * ctx->NMP, h.data(), h.size() * sizeof(double),
                     hipMemcpyHostToDevice));
    return QTP_OK;
}

static int download_disp(qtp_ctx *ctx, double *ref, int b0, int nb)
{
    std::vector<double> h((size_t)nb * 3 * ctx->NMP);
    HIPCHK(hipMemcpy(h.data(), ctx->disp + (size_t)b0 * 3 * ctx->NMP, h.size() * sizeof(double),
                     hipMemcpyDeviceToHost));
    for (int b = 0; b < nb; ++b)
        for (int i = 0; i < ctx->n_mols; ++i)
            for (int d = 0; d < 3; ++d)
                ref[(long)(b0 + b) * 3 * ctx->n_mols + 3 * i + d] = h[((size_t)b * 3 + d) * ctx->NMP + i];
    return QTP_OK;
}

// Caller's half lists (reference layout) -> device half + full lists.  Row
// n_mols-1 is ignored as forces.cpp:489 does.
static int upload_lists(qtp_ctx *ctx, const int *point, const int *list, int b0, int nb,
                        int dev_b0)
{
    const int n = ctx->n_mols;
    std::vector<int> fp((size_t)nb * (n + 1)), fl((size_t)nb * ctx->fcap, 0);
    for (int bb = 0; bb < nb; ++bb) {
        const int b = b0 + bb;
        const int *pt = point + (long)b * n;
        const int *ls = list + (long)b * ctx->list_size;
        std::vector<std::vector<int>> rows(n);
        for (int i = 0; i < n - 1; ++i) {
            const int k0 = pt[i], k1 = pt[i + 1];
            if (k0 < 0 || k1 > ctx->list_size || k1 < k0) {
                ctx->err = "neighbour list: point[] out of range";
                return QTP_ERR_ARG;
            }
            for (int k = k0; k < k1; ++k) {
                const int j = ls[k];
                if (j < 0 || j >= n) {
                    ctx->err = "neighbour list: entry out of range";
                    return QTP_ERR_ARG;
                }
                rows[j].push_back(i);
            }
        }
        long pos = 0;
        for (int i = 0; i < n; ++i) {
            fp[(size_t)bb * (n + 1) + i] = (int)pos;
            std::vector<int> r = rows[i];  // partners j < i (i in row j)
            if (i < n - 1)
                for (int k = pt[i]; k < pt[i + 1]; ++k) r.push_back(ls[k]);
            for (int j : r) {
                if (pos >= ctx->fcap) {
                    ctx->err = "neighbour list: full list exceeds capacity";
                    return QTP_ERR_LIST_OVERFLOW;
                }
                fl[(size_t)bb * ctx->fcap + pos++] = j;
            }
        }
        fp[(size_t)bb * (n + 1) + n] = (int)pos;
    }
    HIPCHK(hipMemcpy(ctx->fpoint + (size_t)dev_b0 * (n + 1), fp.data(), fp.size() * sizeof(int),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->flist + (size_t)dev_b0 * ctx->fcap, fl.data(), fl.size() * sizeof(int),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->point + (size_t)dev_b0 * n, point + (long)b0 * n, (size_t)nb * n * sizeof(int),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->list + (size_t)dev_b0 * ctx->list_size, list + (long)b0 * ctx->list_size,
                     (size_t)nb * ctx->list_size * sizeof(int), hipMemcpyHostToDevice));
    return QTP_OK;
}

static int check_finite(qtp_ctx *ctx, const double *a, long n, const char *what)
{
    for (long i = 0; i < n; ++i)
        if (!std::isfinite(a[i])) {
            ctx->err = std::string(what) + " contains a non-finite value";
            return QTP_ERR_NONFINITE;
        }
    return QTP_OK;
}

static int read_status(qtp_ctx *ctx)
{
    int st = 0;
    HIPCHK(hipMemcpy(&st, ctx->status, sizeof(int), hipMemcpyDeviceToHost));
    if (st & 1) {
        ctx->err = "list size exceeded! (neighbour.cpp:56-58): increase list_size";
        return QTP_ERR_LIST_OVERFLOW;
    }
    if (st & 4) {
        ctx->err = "learned correction: an atom has more neighbours than its descriptor lists "
                   "hold (192 within 6 A, 64 within 4 A)";
        return QTP_ERR_STATE;
    }
    return QTP_OK;
}

static int upload_charges_masses(qtp_ctx *ctx, const double *charge, const double *mass)
{
    if (charge) {
        std::vector<double> c(ctx->NAP, 0.0);
        for (int a = 0; a < ctx->n_atoms; ++a) c[a] = charge[a];
        if (c != ctx->h_charge) {
            ctx->h_charge = c;
            HIPCHK(hipMemcpy(ctx->charge, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice));
            ctx->eself = self_energy(ctx, c.data());
        }
    }
    if (mass) {
        std::vector<double> m(ctx->NAP, 1.0);
        for (int a = 0; a < ctx->n_atoms; ++a) m[a] = mass[3 * a];
        if (m != ctx->h_mass_atom) {
            ctx->h_mass_atom = m;
            std::vector<double> sq(ctx->NAP, 0.0);
            for (int a = 0; a < ctx->n_atoms; ++a) sq[a] = std::sqrt(m[a] / ctx->drv.beta_n);
            HIPCHK(hipMemcpy(ctx->m_atom, m.data(), m.size() * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(ctx->sqm_atom, sq.data(), sq.size() * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    return QTP_OK;
}

static void invalidate_graphs(qtp_ctx *ctx)
{
    for (auto &g : ctx->graph)
        if (g) {
            hipGraphExecDestroy(g);
            g = nullptr;
        }
    for (auto &g : ctx->graph_obs)
        if (g) {
            hipGraphExecDestroy(g);
            g = nullptr;
        }
}

