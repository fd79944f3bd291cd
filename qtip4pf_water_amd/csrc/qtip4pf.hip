// qtip4pf.hip -- host orchestration and C ABI of the gfx950 qTIP4P/F step.
//
// One context = one system (the reference's process globals, water.cpp:29-38)
// with its device buffers, one HIP stream and (resident mode) one captured
// HIP graph per step kind.  See include/qtip4pf.h for the entry points.
// The context and its set-up and launch sequences are in qtp_host_ctx.hpp,
// qtp_host_setup.hpp and qtp_host_run.hpp; this file holds the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "../../include/qtip4pf.h"
#include "qtp_kernels.hpp"
#include "qtp_host.hpp"

using namespace qtp;

// message of the last failed context-free call on this thread (qtp_noalloc.hip)
const char *qtp_noalloc_last_error();

#include "qtp_host_ctx.hpp"
#include "qtp_host_setup.hpp"
#include "qtp_host_nn.hpp"
#include "qtp_host_run.hpp"
#include "qtp_host_tune.hpp"

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int qtp_create(const qtp_params *prm, qtp_ctx **out)
{
    if (!prm || !out) return QTP_ERR_ARG;
    *out = nullptr;
    if (prm->n_mols < 2 || prm->n_beads < 1 || (prm->n_beads != 1 && prm->n_beads % 2 == 1) ||
        prm->boxl[0] <= 0 || prm->boxl[1] <= 0 || prm->boxl[2] <= 0 || prm->eps_ewald <= 0 ||
        prm->eps_ewald >= 1 || prm->T_K <= 0 || prm->dtfs <= 0 || prm->n_beads > 256)
        return QTP_ERR_ARG;
    qtp_ctx *ctx = new qtp_ctx();
    ctx->prm = *prm;
    derive(ctx);
    ctx->P = prm->n_beads;
    ctx->n_mols = prm->n_mols;
    ctx->n_atoms = 3 * prm->n_mols;
    ctx->n_dof = 9 * prm->n_mols;
    ctx->nblk = (ctx->n_atoms + 63) / 64;
    ctx->NAP = ctx->nblk * 64;
    ctx->NMP = ((ctx->n_mols + 63) / 64) * 64;
    ctx->list_size = ctx->drv.list_size;
    ctx->fcap = 2L * ctx->list_size;
    // the real-space erfc approximation covers kappa * r <= QTP_ERFCX_XMAX
    // (derived parameters: kappa * r_c = sqrt(-ln eps_ewald) <= 6.2 for eps >= 2e-17)
    // the largest k_max whose k-space factor tables fit the 160 KiB of LDS
    int kmax_ceiling = 1;
    while (kforce_lds_bytes(kmax_ceiling + 2, KF_WAVES,
                            kforce_chunk(kmax_ceiling + 2, KF_WAVES, 2)) <= 160 * 1024)
        ++kmax_ceiling;
    if (2.0 * ctx->drv.rcut > ctx->drv.min_box * (1.0 + 1e-12) || ctx->drv.k_max < 1 ||
        ctx->drv.k_max > kmax_ceiling || !(ctx->drv.kappa * ctx->drv.rcut <= QTP_ERFCX_XMAX)) {
        ctx->err = "Ewald parameters out of range (r_c must be <= L/2, 1 <= k_max <= " +
                   std::to_string(kmax_ceiling) + ", kappa * r_c <= 6.2); k_max = " +
                   std::to_string(ctx->drv.k_max);
        *out = ctx;
        return QTP_ERR_ARG;
    }
    for (int k = 0; k < 3; ++k) {
        ctx->box.L[k] = prm->boxl[k];
        ctx->box.inv[k] = 1.0 / prm->boxl[k];
    }
    FF &f = ctx->ff;
    f.gamma_vs = RC.gamma_vs;
    f.epslj4 = 4.0 * RC.epslj;
    f.epslj24 = 24.0 * RC.epslj;
    f.sigmalj6 = RC.sigmalj6;
    f.sigmalj12 = RC.sigmalj12;
    {
        // qtip4pfparams.h:106-110 at the context's cutoff (the reference's by default)
        const double rc = ctx->drv.lj_rcut, rs = ctx->drv.r_skin;
        f.e_cut = rc == RC.r_cut ? RC.e_cut
                                 : 4.0 * RC.epslj * (std::pow((RC.sigmalj / rc), 12.0) -
                                                     std::pow((RC.sigmalj / rc), 6.0));
        f.r_cut_sq = rc * rc;
        f.r_skin_sq = rs * rs;
    }
    f.roheq = RC.roheq;
    f.ar2 = RC.ar2;
    f.ar3 = RC.ar3;
    f.ar4fac = RC.ar4fac;
    f.Dr = RC.Dr;
    f.theta_eq = RC.theta_eq;
    f.ktheta = RC.ktheta;
    f.ar2x2 = 2.0 * RC.ar2;
    f.ar3x3 = 3.0 * RC.ar3;
    f.ar4x4 = 4.0 * RC.ar4fac;
    f.half_ktheta = 0.5 * RC.ktheta;
    f.half1mg = 0.5 * (1.0 - RC.gamma_vs);
    ctx->ew.kappa = ctx->drv.kappa;
    ctx->ew.kappa2 = ctx->drv.kappa * ctx->drv.kappa;
    ctx->ew.rcut_sq = ctx->drv.rcut * ctx->drv.rcut;
    ctx->ew.pref2k = 2.0 * ctx->drv.kappa / std::sqrt(RC.pi);
    for (int k = 0; k < 3; ++k) ctx->ew.bk[k] = 2.0 * RC.pi / prm->boxl[k];
    ctx->ew.K = ctx->drv.k_max + 1;

    std::vector<int4> g;
    std::vector<double> cf;
    std::vector<double4> rk;
    build_groups(ctx, g, cf, rk);
    ctx->G = (int)g.size();
    ctx->drv.n_kgroups = ctx->G;
    ctx->ngc = (ctx->G + 255) / 256;
    ctx->ngr = (ctx->G + SR_GPB - 1) / SR_GPB;
    const int K = ctx->ew.K;
    // sites per structure-factor chunk: 48 KiB of factor tables, a multiple
    // of 4 (the matrix-core k-steps) below 64
    ctx->AC = std::max(4, std::min(64, (int)(48 * 1024 / (3 * K * 16))) / 4 * 4);
    ctx->nac = (ctx->n_atoms + ctx->AC - 1) / ctx->AC;
    ctx->ksum_lds = (size_t)ctx->AC * 3 * K * sizeof(double2);
    // two 12-wave blocks or three 6-wave (split) blocks per CU
    ctx->kf_chunk = kforce_chunk(K, KF_WAVES, 2);
    ctx->kf_chunk_split = kforce_chunk(K, KF_SPLIT_WAVES, 3);
    ctx->kforce_lds = kforce_lds_bytes(K, KF_WAVES, ctx->kf_chunk);
    ctx->kf_split_ok = (long)ctx->P * ((ctx->n_atoms + KF_SITES - 1) / KF_SITES) <= 512;
    ctx->kf_split = ctx->kf_split_ok;
    ctx->kforce_lds_split = kforce_lds_bytes(K, KF_SPLIT_WAVES, ctx->kf_chunk_split);
    std::vector<int> gcol;
    std::vector<int4> mf_tiles, mf_rt, mf_rt_small, ks_col;
    std::vector<int2> ks_unit;
    {
        const char *kv = std::getenv("QTP_KSPACE");
        const bool want_mf = !(kv && std::strcmp(kv, "vector") == 0);
        ctx->kf_mf_ok = 2 * K <= 4 * MF_MAXKS && ctx->G > 0 && ctx->AC % 4 == 0 &&
                        build_mf_layout(g, gcol, mf_tiles, ctx->mf_Mp, ctx->mf_nks, ks_col, ks_unit);
        ctx->kf_mf = ctx->kf_mf_ok && want_mf;
        if (ctx->kf_mf_ok) {
            mf_rt = mf_tile_order(mf_tiles, MF_WAVES);
            mf_rt_small = mf_tile_order(mf_tiles, MF_WAVES_SMALL);
            ctx->mf_nrt = (int)mf_rt.size();
            ctx->mf_nrt_small = (int)mf_rt_small.size();
            ctx->ks_nunit = (int)ks_unit.size();
            ctx->kforce_mf_lds = kforce_mf_lds_bytes(K, ctx->mf_nks, MF_WAVES);
            ctx->kforce_mf_lds_small = kforce_mf_lds_bytes(K, ctx->mf_nks, MF_WAVES_SMALL);
            ctx->ks_ncol4 = (int)ks_col.size();
            ctx->ksum_mf_lds = ksum_mf_lds_bytes(K, ctx->AC, ctx->ks_ncol4, ctx->ks_nunit);
        }
        if (ctx->kf_mf) ctx->kf_split = false;
    }
    std::vector<int2> tl;
    for (int I = 0; I < ctx->nblk; ++I)
        for (int J = I; J < ctx->nblk; ++J) tl.push_back(make_int2(I, J));
    ctx->ntile = (int)tl.size();
    ctx->rs_cull = std::min(ctx->box.L[0], std::min(ctx->box.L[1], ctx->box.L[2])) >=
                   RS_CULL_MIN_BOX * std::sqrt(ctx->ew.rcut_sq);
    // energy partials sized for the widest layout (a wave per molecule)
    set_lj_lanes(ctx, 64);
    ctx->ep_cap = ctx->ep_stride;
    // a whole wave per molecule when one launch covers few molecules
    set_lj_lanes(ctx, (long)std::max(ctx->P, 1) * ctx->n_mols <= 2048 ? 64 : 16);
    ctx->ptile = (3 * ctx->NAP + 63) / 64;
    ctx->prop_lds = 0;
    ctx->cells_ok = cell_grid(ctx->drv.r_skin, ctx->box.L, ctx->n_mols, ctx->cg);
    *out = ctx;

    // a blocking stream: host-synchronous hipMemcpy/hipMemset on the null
    // stream are ordered with this stream's kernels
    if (hipStreamCreate(&ctx->stream) != hipSuccess || hipStreamCreate(&ctx->s2) != hipSuccess ||
        hipStreamCreate(&ctx->s3) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_nn_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_nn_join, hipEventDisableTiming) != hipSuccess) {
        ctx->err = "hipStreamCreate failed";
        return QTP_ERR_HIP;
    }
    // matrix-core k-space kernels with more than the default 64 KiB of LDS
    // (large k-sets); without the attribute they are not used
    if (ctx->kf_mf_ok) {
        const bool hi = ctx->mf_nks > MF_KS_LO;
        bool ok = ctx->ksum_mf_lds <= 160 * 1024 && ctx->kforce_mf_lds <= 160 * 1024;
        if (ok && ctx->ksum_mf_lds > 64 * 1024)
            ok = hipFuncSetAttribute((const void *)k_ksum_mf,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)ctx->ksum_mf_lds) == hipSuccess;
        if (ok && ctx->kforce_mf_lds > 64 * 1024)
            ok = hipFuncSetAttribute(hi ? (const void *)k_kforce_mf<MF_WAVES, MF_MAXKS>
                                        : (const void *)k_kforce_mf<MF_WAVES, MF_KS_LO>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)ctx->kforce_mf_lds) == hipSuccess;
        if (!ok) {
            ctx->kf_mf_ok = false;
            ctx->kf_mf = false;
            ctx->kf_split = ctx->kf_split_ok;
        }
    }
    if (ctx->kforce_lds > 64 * 1024 || ctx->ksum_lds > 64 * 1024) {
        if (ctx->kforce_lds > 160 * 1024 ||
            hipFuncSetAttribute((const void *)k_kforce<KF_WAVES, false>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)ctx->kforce_lds) != hipSuccess ||
            (ctx->kforce_lds_split > 64 * 1024 &&
             hipFuncSetAttribute((const void *)k_kforce<KF_SPLIT_WAVES, true>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)ctx->kforce_lds_split) != hipSuccess)) {
            ctx->err = "LDS request above the device limit (k_max too large)";
            return QTP_ERR_ARG;
        }
    }
    // the generic bead transform keeps P x 64 doubles in LDS (128 KiB at P = 256)
    if (nm_generic_lds(ctx->P) > 64 * 1024 &&
        hipFuncSetAttribute((const void *)k_nm_generic, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)nm_generic_lds(ctx->P)) != hipSuccess) {
        ctx->err = "LDS request of the bead transform above the device limit";
        return QTP_ERR_ARG;
    }
    int st = QTP_OK;
    if ((st = dalloc(ctx, &ctx->groups, ctx->G)) || (st = dalloc(ctx, &ctx->gcf, ctx->G)) ||
        (st = dalloc(ctx, &ctx->grk, ctx->G)) || (st = dalloc(ctx, &ctx->tiles, ctx->ntile)) ||
        (st = dalloc(ctx, &ctx->Cm, (size_t)ctx->P * ctx->P)) ||
        (st = dalloc(ctx, &ctx->mode, ctx->P)) || (st = dalloc(ctx, &ctx->ms, ctx->P)) ||
        (ctx->kf_mf_ok && ((st = dalloc(ctx, &ctx->gcol, ctx->G)) ||
                        (st = dalloc(ctx, &ctx->mf_rt, ctx->mf_nrt)) ||
                        (st = dalloc(ctx, &ctx->mf_rt_small, ctx->mf_nrt_small)) ||
                        (st = dalloc(ctx, &ctx->ks_col, ks_col.size())) ||
                        (st = dalloc(ctx, &ctx->ks_unit, ks_unit.size())))) ||
        (st = alloc_all(ctx)))
        return st;
    if (ctx->kf_mf_ok &&
        (hipMemcpy(ctx->gcol, gcol.data(), gcol.size() * sizeof(int), hipMemcpyHostToDevice) ||
         hipMemcpy(ctx->mf_rt, mf_rt.data(), mf_rt.size() * sizeof(int4), hipMemcpyHostToDevice) ||
         hipMemcpy(ctx->mf_rt_small, mf_rt_small.data(), mf_rt_small.size() * sizeof(int4),
                   hipMemcpyHostToDevice) ||
         hipMemcpy(ctx->ks_col, ks_col.data(), ks_col.size() * sizeof(int4), hipMemcpyHostToDevice) ||
         hipMemcpy(ctx->ks_unit, ks_unit.data(), ks_unit.size() * sizeof(int2), hipMemcpyHostToDevice))) {
        ctx->err = "constant upload failed";
        return QTP_ERR_HIP;
    }
    if (hipMemcpy(ctx->groups, g.data(), g.size() * sizeof(int4), hipMemcpyHostToDevice) ||
        hipMemcpy(ctx->gcf, cf.data(), cf.size() * sizeof(double), hipMemcpyHostToDevice) ||
        hipMemcpy(ctx->grk, rk.data(), rk.size() * sizeof(double4), hipMemcpyHostToDevice) ||
        hipMemcpy(ctx->tiles, tl.data(), tl.size() * sizeof(int2), hipMemcpyHostToDevice)) {
        ctx->err = "constant upload failed";
        return QTP_ERR_HIP;
    }
    CHK(upload_constants(ctx));
    HIPCHK(hipDeviceSynchronize());
    return QTP_OK;
}

void qtp_destroy(qtp_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    invalidate_graphs(ctx);
    for (auto e : ctx->evpool) hipEventDestroy(e);
    void *ptrs[] = {ctx->q, ctx->p, ctx->F, ctx->fph, ctx->disp, ctx->m_atom, ctx->sqm_atom,
                    ctx->charge, ctx->point, ctx->list, ctx->fpoint, ctx->flist, ctx->cnt,
                    ctx->flag, ctx->status, ctx->nreb, ctx->qpm, ctx->vpart, ctx->vout, ctx->rpart, ctx->rs_sites, ctx->rs_box, ctx->rs_far, ctx->nl_bm, ctx->cell_of, ctx->crank, ctx->cstart, ctx->member, ctx->cell_pos,
                    ctx->spart, ctx->sprime, ctx->flj, ctx->fspl, ctx->spl_coef, ctx->fpm, ctx->fk2, ctx->wm, ctx->gcol, ctx->mf_rt, ctx->mf_rt_small, ctx->ks_col, ctx->ks_unit, ctx->Ve, ctx->epart,
                    ctx->scal, ctx->out, ctx->heat, ctx->heat_step, ctx->step_ctr, ctx->groups,
                    ctx->gcf, ctx->grk, ctx->tiles, ctx->Cm, ctx->mode, ctx->nq, ctx->np,
                    ctx->ppart, ctx->hpart, ctx->dev, ctx->obs, ctx->obs_out, ctx->ntrace, ctx->trace,
                    ctx->fnn, ctx->ms, ctx->cmf_acc, ctx->cmf_n, ctx->qc, ctx->rdf_oo,
                    ctx->rdf_oh};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (ctx->s2) hipStreamSynchronize(ctx->s2);
    if (ctx->s3) hipStreamSynchronize(ctx->s3);
    if (ctx->ev_fork) hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) hipEventDestroy(ctx->ev_join);
    if (ctx->ev_nn_fork) hipEventDestroy(ctx->ev_nn_fork);
    if (ctx->ev_nn_join) hipEventDestroy(ctx->ev_nn_join);
    if (ctx->s2) hipStreamDestroy(ctx->s2);
    if (ctx->s3) hipStreamDestroy(ctx->s3);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *qtp_last_error(const qtp_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : qtp_noalloc_last_error();
}

int qtp_get_derived(const qtp_ctx *ctx, qtp_derived *out)
{
    if (!ctx || !out) return QTP_ERR_ARG;
    *out = ctx->drv;
    return QTP_OK;
}

int qtp_get_stream(const qtp_ctx *ctx, void **stream)
{
    if (!ctx || !stream) return QTP_ERR_ARG;
    *stream = (void *)ctx->stream;
    return QTP_OK;
}

// ---- reference-shaped calls --------------------------------------------------

static int get_forces_impl(qtp_ctx *ctx, const double *q, double *F, const double *charge,
                            const int *point, const int *list, int bead, double *energy)
{
    if (!ctx || !q || !F || !point || !list || bead < 0 || bead >= ctx->P) return QTP_ERR_ARG;
    CHK(check_finite(ctx, q + (long)bead * ctx->n_dof, ctx->n_dof, "positions"));
    CHK(upload_charges_masses(ctx, charge, nullptr));
    // stage the bead's positions in evaluation slot 0 of qpm, its lists in slot `bead`
    CHK(upload_planar(ctx, ctx->qpm, q + (long)bead * ctx->n_dof, 1));
    CHK(upload_lists(ctx, point, list, bead, 1, bead));
    CHK(run_phase(ctx, ctx->qpm, 1, 0, bead, ctx->fpm, ctx->Ve));
    double e = 0.0;
    HIPCHK(hipMemcpyAsync(&e, ctx->Ve, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CHK(download_planar(ctx, ctx->fpm, 1, F + (long)bead * ctx->n_dof));
    CHK(flush_timing(ctx));
    if (energy) *energy = e;
    ctx->state_set = false;
    return QTP_OK;
}

int qtp_get_forces_virial(qtp_ctx *ctx, const double *q, double *F, const double *charge,
                          const int *point, const int *list, int bead, double *energy,
                          double virial[6])
{
    if (!ctx || !q || !F || !point || !list || !virial || bead < 0 || bead >= ctx->P)
        return QTP_ERR_ARG;
    SplineScope sp(ctx, false);
    CHK(check_finite(ctx, q + (long)bead * ctx->n_dof, ctx->n_dof, "positions"));
    CHK(upload_charges_masses(ctx, charge, nullptr));
    CHK(upload_planar(ctx, ctx->qpm, q + (long)bead * ctx->n_dof, 1));
    CHK(upload_lists(ctx, point, list, bead, 1, bead));
    CHK(run_phase(ctx, ctx->qpm, 1, 0, bead, ctx->fpm, ctx->Ve, ctx->vout));
    double e = 0.0;
    HIPCHK(hipMemcpyAsync(&e, ctx->Ve, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(virial, ctx->vout, 6 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CHK(download_planar(ctx, ctx->fpm, 1, F + (long)bead * ctx->n_dof));
    CHK(flush_timing(ctx));
    if (energy) *energy = e;
    ctx->state_set = false;
    return QTP_OK;
}

static int force_eval_sc_impl(qtp_ctx *ctx, const double *q, double *F_sc, const double *charge,
                               const double *mass, const int *point, const int *list,
                               double *sc_pot, double *fphys, double *energy)
{
    if (!ctx || !q || !F_sc || !point || !list || ctx->P < 2) return QTP_ERR_ARG;
    const int P = ctx->P;
    CHK(check_finite(ctx, q, (long)P * ctx->n_dof, "positions"));
    CHK(upload_charges_masses(ctx, charge, mass));
    CHK(upload_planar(ctx, ctx->q, q, P));
    CHK(upload_lists(ctx, point, list, 0, P, 0));
    CHK(run_sc_forces(ctx, 0, nullptr));
    double o[2];
    HIPCHK(hipMemcpyAsync(o, ctx->out, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CHK(download_planar(ctx, ctx->F, P, F_sc));
    if (fphys) CHK(download_planar(ctx, ctx->fph, P, fphys));
    CHK(flush_timing(ctx));
    if (energy) *energy = o[0];
    if (sc_pot) *sc_pot = o[1];
    ctx->state_set = false;
    return QTP_OK;
}

int qtp_neighbour_list(qtp_ctx *ctx, const double *q, int *point, int *list, double *disp,
                       int bead)
{
    if (!ctx || !q || !point || !list || bead < 0 || bead >= ctx->P) return QTP_ERR_ARG;
    CHK(check_finite(ctx, q + (long)bead * ctx->n_dof, ctx->n_dof, "positions"));
    const long N3 = 3L * ctx->NAP;
    CHK(upload_planar(ctx, ctx->q + bead * N3, q + (long)bead * ctx->n_dof, 1));
    std::vector<int> fl(ctx->P, 0);
    fl[bead] = 1;
    HIPCHK(hipMemcpy(ctx->flag, fl.data(), ctx->P * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(ctx->status, 0, sizeof(int)));
    const int nrb = (ctx->n_mols + 3) / 4;
    hipStream_t s = ctx->stream;
    if (use_cells(ctx)) {
        CHK(run_cell_lists(ctx, 0, 0, nullptr));
    } else {
        TimeScope ts(ctx, "nlist");
        k_nl_count<<<ctx->P * nrb, 256, 0, s>>>(ctx->n_mols, ctx->NAP, ctx->q, ctx->box,
                                                ctx->ff.r_skin_sq, ctx->flag, ctx->cnt);
        k_nl_scan<<<ctx->P, 1024, 0, s>>>(ctx->n_mols, ctx->list_size, ctx->fcap, ctx->flag,
                                          ctx->cnt, ctx->point, ctx->fpoint, ctx->list, ctx->disp,
                                          ctx->NMP, ctx->status);
        k_nl_fill<<<ctx->P * nrb, 256, 0, s>>>(ctx->n_mols, ctx->NAP, ctx->q, ctx->box,
                                               ctx->ff.r_skin_sq, ctx->list_size, ctx->fcap, ctx->flag,
                                               ctx->point, ctx->fpoint, ctx->list, ctx->flist);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    CHK(flush_timing(ctx));
    int st = read_status(ctx);
    HIPCHK(hipMemcpy(point + (long)bead * ctx->n_mols, ctx->point + (size_t)bead * ctx->n_mols,
                     ctx->n_mols * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(list + (long)bead * ctx->list_size, ctx->list + (size_t)bead * ctx->list_size,
                     ctx->list_size * sizeof(int), hipMemcpyDeviceToHost));
    if (disp)
        for (int i = 0; i < 3 * ctx->n_mols; ++i) disp[(long)bead * 3 * ctx->n_mols + i] = 0.0;
    ctx->state_set = false;
    return st;
}

int qtp_ftransform(qtp_ctx *ctx, double *a, int length, int is_forwards)
{
    if (!ctx || !a || length < 1) return QTP_ERR_ARG;
    const int P = ctx->P;
    const size_t n = (size_t)P * length;
    double *d = ctx->nq;  // scratch of P * 3 * NAP doubles
    bool own = false;
    if (n > (size_t)P * 3 * ctx->NAP) {
        HIPCHK(hipMalloc((void **)&d, n * sizeof(double)));
        own = true;
    }
    HIPCHK(hipMemcpyAsync(d, a, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const size_t lds = nm_generic_lds(P);
    {
        TimeScope ts(ctx, "nm");
        k_nm_generic<<<(length + 63) / 64, 256, lds, ctx->stream>>>(P, length, ctx->Cm, d,
                                                                   is_forwards ? 1 : 0);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(a, d, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (own) HIPCHK(hipFree(d));
    ctx->state_set = false;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    CHK(flush_timing(ctx));
    return QTP_OK;
}

int qtp_pile_g(qtp_ctx *ctx, double *p_nm, const double *mass, const double *dev, double *heat)
{
    if (!ctx || !p_nm || !dev || ctx->P < 2) return QTP_ERR_ARG;
    const int P = ctx->P;
    CHK(upload_charges_masses(ctx, nullptr, mass));
    CHK(upload_planar(ctx, ctx->np, p_nm, P));
    HIPCHK(hipMemcpy(ctx->dev, dev, ctx->dev_cap * sizeof(double), hipMemcpyHostToDevice));
    const int nt = (3 * ctx->NAP + 255) / 256;
    {
        TimeScope ts(ctx, "pile");
        k_pile_a<<<nt, 256, 0, ctx->stream>>>(ctx->n_atoms, ctx->NAP, nt, ctx->np, ctx->m_atom,
                                              ctx->ppart);
        k_pile_b<<<nt, 256, 0, ctx->stream>>>(P, ctx->n_atoms, ctx->NAP, ctx->n_dof, nt,
                                              ctx->drv.beta_n, ctx->mode, ctx->m_atom,
                                              ctx->sqm_atom, ctx->ppart, ctx->dev, ctx->np,
                                              ctx->hpart);
        k_heat<<<1, 1024, 0, ctx->stream>>>(nt, nt, ctx->ppart, ctx->hpart, nullptr, ctx->heat_step);
    }
    HIPCHK(hipGetLastError());
    double h = 0.0;
    HIPCHK(hipMemcpyAsync(&h, ctx->heat_step, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CHK(download_planar(ctx, ctx->np, P, p_nm));
    CHK(flush_timing(ctx));
    if (heat) *heat = h;
    ctx->state_set = false;
    return QTP_OK;
}

static int update_sc_impl(qtp_ctx *ctx, double *q, double *p, const double *mass, double *F,
                           double *fphys, const double *charge, int *point, int *list,
                           double *disp, const double *dev, int is_nvt, double *heat,
                           double *sc_pot, double *energy)
{
    if (!ctx || !q || !p || !F || !point || !list || !disp || ctx->P < 2 || (is_nvt && !dev))
        return QTP_ERR_ARG;
    if (ctx->pimd_mode != QTP_MODE_PIMD) {
        ctx->err = "update_sc_nospline steps plain PIMD (qtp_set_pimd_mode QTP_MODE_PIMD)";
        return QTP_ERR_STATE;
    }
    const int P = ctx->P;
    CHK(check_finite(ctx, q, (long)P * ctx->n_dof, "positions"));
    CHK(upload_charges_masses(ctx, charge, mass));
    CHK(upload_planar(ctx, ctx->q, q, P));
    CHK(upload_planar(ctx, ctx->p, p, P));
    CHK(upload_planar(ctx, ctx->F, F, P));
    CHK(upload_disp(ctx, disp, 0, P));
    CHK(upload_lists(ctx, point, list, 0, P, 0));
    if (is_nvt)
        HIPCHK(hipMemcpy(ctx->dev, dev, ctx->dev_cap * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(ctx->status, 0, sizeof(int)));
    double h0 = heat ? *heat : 0.0;
    HIPCHK(hipMemcpy(ctx->heat, &h0, sizeof(double), hipMemcpyHostToDevice));
    CHK(run_sc_step(ctx, is_nvt, 1, nullptr));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    CHK(flush_timing(ctx));
    CHK(read_status(ctx));
    double o[2];
    HIPCHK(hipMemcpy(o, ctx->out, 2 * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&h0, ctx->heat, sizeof(double), hipMemcpyDeviceToHost));
    CHK(download_planar(ctx, ctx->q, P, q));
    CHK(download_planar(ctx, ctx->p, P, p));
    CHK(download_planar(ctx, ctx->F, P, F));
    if (fphys) CHK(download_planar(ctx, ctx->fph, P, fphys));
    CHK(download_disp(ctx, disp, 0, P));
    std::vector<int> fl(P);
    HIPCHK(hipMemcpy(fl.data(), ctx->flag, P * sizeof(int), hipMemcpyDeviceToHost));
    for (int b = 0; b < P; ++b)
        if (fl[b]) {
            HIPCHK(hipMemcpy(point + (long)b * ctx->n_mols, ctx->point + (size_t)b * ctx->n_mols,
                             ctx->n_mols * sizeof(int), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(list + (long)b * ctx->list_size, ctx->list + (size_t)b * ctx->list_size,
                             ctx->list_size * sizeof(int), hipMemcpyDeviceToHost));
        }
    if (heat) *heat = h0;
    if (energy) *energy = o[0];
    if (sc_pot) *sc_pot = o[1];
    ctx->state_set = false;
    return QTP_OK;
}

static int update_class_impl(qtp_ctx *ctx, double *q, double *p, const double *mass,
                                         double *F, const double *charge, int *point, int *list,
                                         double *disp, double *energy)
{
    if (!ctx || !q || !p || !F || !point || !list || !disp || ctx->P != 1) return QTP_ERR_ARG;
    CHK(check_finite(ctx, q, ctx->n_dof, "positions"));
    CHK(upload_charges_masses(ctx, charge, mass));
    CHK(upload_planar(ctx, ctx->q, q, 1));
    CHK(upload_planar(ctx, ctx->p, p, 1));
    CHK(upload_planar(ctx, ctx->F, F, 1));
    CHK(upload_disp(ctx, disp, 0, 1));
    CHK(upload_lists(ctx, point, list, 0, 1, 0));
    HIPCHK(hipMemset(ctx->status, 0, sizeof(int)));
    CHK(run_class_step(ctx, nullptr));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    CHK(flush_timing(ctx));
    CHK(read_status(ctx));
    double o[2];
    HIPCHK(hipMemcpy(o, ctx->out, 2 * sizeof(double), hipMemcpyDeviceToHost));
    CHK(download_planar(ctx, ctx->q, 1, q));
    CHK(download_planar(ctx, ctx->p, 1, p));
    CHK(download_planar(ctx, ctx->F, 1, F));
    CHK(download_disp(ctx, disp, 0, 1));
    int fl = 0;
    HIPCHK(hipMemcpy(&fl, ctx->flag, sizeof(int), hipMemcpyDeviceToHost));
    if (fl) {
        HIPCHK(hipMemcpy(point, ctx->point, ctx->n_mols * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(list, ctx->list, ctx->list_size * sizeof(int), hipMemcpyDeviceToHost));
    }
    if (energy) *energy = o[0];
    ctx->state_set = false;
    return QTP_OK;
}

int qtp_get_kcv(qtp_ctx *ctx, const double *q, const double *fphys, double *kcv)
{
    if (!ctx || !q || !fphys || !kcv) return QTP_ERR_ARG;
    const int P = ctx->P;
    CHK(upload_planar(ctx, ctx->nq, q, P));
    CHK(upload_planar(ctx, ctx->np, fphys, P));
    const int nb = (3 * ctx->NAP + 255) / 256;
    k_kcv<<<nb, 256, 0, ctx->stream>>>(P, ctx->n_atoms, ctx->NAP, ctx->nq, ctx->np, ctx->obs);
    HIPCHK(hipGetLastError());
    std::vector<double> part(nb);
    HIPCHK(hipMemcpyAsync(part.data(), ctx->obs, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double tq = 0.0;
    for (double v : part) tq += v;
    tq /= P;
    *kcv = 0.5 * ctx->n_dof / ctx->drv.beta + tq;
    ctx->state_set = false;
    return QTP_OK;
}

int qtp_get_conserved(qtp_ctx *ctx, const double *q, const double *p, const double *mass,
                      double energy, double *cons)
{
    if (!ctx || !q || !p || !cons) return QTP_ERR_ARG;
    const int P = ctx->P;
    CHK(upload_charges_masses(ctx, nullptr, mass));
    CHK(upload_planar(ctx, ctx->nq, q, P));
    CHK(upload_planar(ctx, ctx->np, p, P));
    const int len = 3 * ctx->NAP;
    const size_t lds = nm_generic_lds(P);
    k_nm_generic<<<(len + 63) / 64, 256, lds, ctx->stream>>>(P, len, ctx->Cm, ctx->nq, 1);
    k_nm_generic<<<(len + 63) / 64, 256, lds, ctx->stream>>>(P, len, ctx->Cm, ctx->np, 1);
    const int nb = (3 * ctx->NAP + 255) / 256;
    k_cons<<<nb, 256, 0, ctx->stream>>>(P, ctx->n_atoms, ctx->NAP, ctx->nq, ctx->np, ctx->mode,
                                        ctx->m_atom, ctx->obs);
    HIPCHK(hipGetLastError());
    std::vector<double> part(nb);
    HIPCHK(hipMemcpyAsync(part.data(), ctx->obs, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (double v : part) s += v;
    *cons = energy + s;
    ctx->state_set = false;
    return QTP_OK;
}

// ---- splined pair potentials (fast-CMD path) ------------------------------------

// Uniform cubic B-spline interpolation of f[0..n-1] at a + j h, clamped with
// endpoint slopes from the 5-point one-sided differences (O(h^4)): the
// construction of Boost.Math's cubic_b_spline, which the reference builds
// from its spline files (simulation.cpp:60, 84, 111).  Rows
// c_{j-1} + 4 c_j + c_{j+1} = 6 f_j, c_{-1} = c_1 - 2h s'(a),
// c_n = c_{n-2} + 2h s'(b); tridiagonal solve in long double.
static void spline_fit(const double *f, int n, double h, double *c)
{
    const long double hl = h;
    const long double sa = (-25.0L * f[0] + 48.0L * f[1] - 36.0L * f[2] + 16.0L * f[3] -
                            3.0L * f[4]) / (12.0L * hl);
    const long double sb = (25.0L * f[n - 1] - 48.0L * f[n - 2] + 36.0L * f[n - 3] -
                            16.0L * f[n - 4] + 3.0L * f[n - 5]) / (12.0L * hl);
    std::vector<long double> dg(n, 4.0L), up(n, 1.0L), lo(n, 1.0L), r(n), x(n);
    for (int j = 0; j < n; ++j) r[j] = 6.0L * f[j];
    up[0] = 2.0L;
    r[0] += 2.0L * hl * sa;
    lo[n - 1] = 2.0L;
    r[n - 1] -= 2.0L * hl * sb;
    for (int j = 1; j < n; ++j) {
        const long double m = lo[j] / dg[j - 1];
        dg[j] -= m * up[j - 1];
        r[j] -= m * r[j - 1];
    }
    x[n - 1] = r[n - 1] / dg[n - 1];
    for (int j = n - 2; j >= 0; --j) x[j] = (r[j] - up[j] * x[j + 1]) / dg[j];
    for (int j = 0; j < n; ++j) c[j + 1] = (double)x[j];
    c[0] = (double)(x[1] - 2.0L * hl * sa);
    c[n + 1] = (double)(x[n - 2] + 2.0L * hl * sb);
}

int qtp_set_splines(qtp_ctx *ctx, const double *oo, int n_oo, const double *oh, int n_oh,
                    const double *hh, int n_hh, double start, double step)
{
    if (!ctx || !oo || !oh || !hh || n_oo < 5 || n_oh < 5 || n_hh < 5 || !(step > 0.0) ||
        !std::isfinite(start))
        return QTP_ERR_ARG;
    const double *tab[3] = {oo, oh, hh};
    const int nn[3] = {n_oo, n_oh, n_hh};
    for (int t = 0; t < 3; ++t) CHK(check_finite(ctx, tab[t], nn[t], "spline table"));
    std::vector<double> coef;
    SplTab st{};
    st.a = start;
    st.h = step;
    st.inv_h = 1.0 / step;
    for (int t = 0; t < 3; ++t) {
        st.n[t] = nn[t];
        st.off[t] = (int)coef.size();
        coef.resize(coef.size() + nn[t] + 2);
        spline_fit(tab[t], nn[t], step, coef.data() + st.off[t]);
        // simulation.cpp:57-58: the table ends at start + step (n - 1)
        const double cut = start + step * (nn[t] - 1);
        st.cutsq[t] = cut * cut;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->spl_coef) hipFree(ctx->spl_coef);
    ctx->spl_coef = nullptr;
    CHK(dalloc(ctx, &ctx->spl_coef, coef.size()));
    HIPCHK(hipMemcpy(ctx->spl_coef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice));
    ctx->spl = st;
    ctx->spl_loaded = true;
    ctx->spl_resident = true;
    invalidate_graphs(ctx);
    ctx->state_set = false;
    return QTP_OK;
}

int qtp_clear_splines(qtp_ctx *ctx)
{
    if (!ctx) return QTP_ERR_ARG;
    ctx->spl_resident = false;
    invalidate_graphs(ctx);
    ctx->state_set = false;
    return QTP_OK;
}

static int need_splines(qtp_ctx *ctx)
{
    if (!ctx) return QTP_ERR_ARG;
    if (!ctx->spl_loaded) {
        ctx->err = "no spline tables loaded (qtp_set_splines)";
        return QTP_ERR_STATE;
    }
    return QTP_OK;
}

int qtp_get_forces_nospline(qtp_ctx *ctx, const double *q, double *F, const double *charge,
                            const int *point, const int *list, int bead, double *energy)
{
    if (!ctx) return QTP_ERR_ARG;
    SplineScope sp(ctx, false);
    return get_forces_impl(ctx, q, F, charge, point, list, bead, energy);
}

int qtp_get_forces_wspline(qtp_ctx *ctx, const double *q, double *F, const double *charge,
                           const int *point, const int *list, int bead, double *energy)
{
    CHK(need_splines(ctx));
    SplineScope sp(ctx, true);
    return get_forces_impl(ctx, q, F, charge, point, list, bead, energy);
}

int qtp_force_eval_sc_nospline(qtp_ctx *ctx, const double *q, double *F_sc, const double *charge,
                               const double *mass, const int *point, const int *list,
                               double *sc_pot, double *fphys, double *energy)
{
    if (!ctx) return QTP_ERR_ARG;
    SplineScope sp(ctx, false);
    return force_eval_sc_impl(ctx, q, F_sc, charge, mass, point, list, sc_pot, fphys, energy);
}

int qtp_force_eval_sc_wspline(qtp_ctx *ctx, const double *q, double *F_sc, const double *charge,
                              const double *mass, const int *point, const int *list,
                              double *sc_pot, double *fphys, double *energy)
{
    CHK(need_splines(ctx));
    SplineScope sp(ctx, true);
    return force_eval_sc_impl(ctx, q, F_sc, charge, mass, point, list, sc_pot, fphys, energy);
}

int qtp_update_sc_nospline(qtp_ctx *ctx, double *q, double *p, const double *mass, double *F,
                           double *fphys, const double *charge, int *point, int *list,
                           double *disp, const double *dev, int is_nvt, double *heat,
                           double *sc_pot, double *energy)
{
    if (!ctx) return QTP_ERR_ARG;
    SplineScope sp(ctx, false);
    return update_sc_impl(ctx, q, p, mass, F, fphys, charge, point, list, disp, dev, is_nvt, heat,
                          sc_pot, energy);
}

int qtp_update_sc_wspline(qtp_ctx *ctx, double *q, double *p, const double *mass, double *F,
                          double *fphys, const double *charge, int *point, int *list,
                          double *disp, const double *dev, int is_nvt, double *heat,
                          double *sc_pot, double *energy)
{
    CHK(need_splines(ctx));
    SplineScope sp(ctx, true);
    return update_sc_impl(ctx, q, p, mass, F, fphys, charge, point, list, disp, dev, is_nvt, heat,
                          sc_pot, energy);
}

int qtp_update_trajectory_class_nospline(qtp_ctx *ctx, double *q, double *p, const double *mass,
                                         double *F, const double *charge, int *point, int *list,
                                         double *disp, double *energy)
{
    if (!ctx) return QTP_ERR_ARG;
    SplineScope sp(ctx, false);
    return update_class_impl(ctx, q, p, mass, F, charge, point, list, disp, energy);
}

int qtp_update_trajectory_class_wspline(qtp_ctx *ctx, double *q, double *p, const double *mass,
                                        double *F, const double *charge, int *point, int *list,
                                        double *disp, double *energy)
{
    CHK(need_splines(ctx));
    SplineScope sp(ctx, true);
    return update_class_impl(ctx, q, p, mass, F, charge, point, list, disp, energy);
}

// ---- host rng --------------------------------------------------------------

int qtp_host_rng_create(uint32_t seed, qtp_host_rng **out)
{
    if (!out) return QTP_ERR_ARG;
    *out = new qtp_host_rng();
    (*out)->gen.seed(seed);
    return QTP_OK;
}

void qtp_host_rng_destroy(qtp_host_rng *r) { delete r; }

// thermostat.cpp:71, 88-98, 124: one normal_distribution object for R, R_add
// and every xi; a separate gamma_distribution for the centroid.
int qtp_host_rng_draw_pile(qtp_host_rng *r, int n_atoms, int n_beads, double *dev)
{
    if (!r || !dev || n_atoms < 1 || n_beads < 1) return QTP_ERR_ARG;
    std::normal_distribution<double> gauss_dev(0.0, 1.0);
    const int dof = 3 * n_atoms;
    dev[0] = gauss_dev(r->gen);
    if (dof % 2 == 0) {
        std::gamma_distribution<double> gamma_dev((dof - 2) / 2, 1.0);
        dev[1] = gauss_dev(r->gen);
        dev[2] = gamma_dev(r->gen);
    } else {
        std::gamma_distribution<double> gamma_dev((dof - 1) / 2, 1.0);
        dev[1] = 0.0;
        dev[2] = gamma_dev(r->gen);
    }
    const long n = (long)(n_beads - 1) * dof;
    for (long i = 0; i < n; ++i) dev[3 + i] = gauss_dev(r->gen);
    return QTP_OK;
}

// ---- resident mode ------------------------------------------------------------

int qtp_set_state(qtp_ctx *ctx, const double *q, const double *p, const double *charge,
                  const double *mass)
{
    if (!ctx || !q || !p) return QTP_ERR_ARG;
    SplineScope sp(ctx, ctx->spl_resident);
    const int P = ctx->P;
    CHK(check_finite(ctx, q, (long)P * ctx->n_dof, "positions"));
    CHK(check_finite(ctx, p, (long)P * ctx->n_dof, "momenta"));
    invalidate_graphs(ctx);
    CHK(upload_charges_masses(ctx, charge, mass));
    CHK(upload_planar(ctx, ctx->q, q, P));
    CHK(upload_planar(ctx, ctx->p, p, P));
    HIPCHK(hipMemset(ctx->disp, 0, (size_t)P * 3 * ctx->NMP * sizeof(double)));
    HIPCHK(hipMemset(ctx->status, 0, sizeof(int)));
    HIPCHK(hipMemset(ctx->heat, 0, sizeof(double)));
    HIPCHK(hipMemset(ctx->nreb, 0, P * sizeof(double)));
    HIPCHK(hipMemset(ctx->step_ctr, 0, sizeof(unsigned long long)));
    CHK(run_lists(ctx, 1));
    if (P >= 2) {
        CHK(run_sc_forces(ctx, 0, nullptr));
    } else {
        CHK(class_initial_forces(ctx));
    }
    HIPCHK(hipMemsetAsync(ctx->nreb, 0, P * sizeof(double), ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    CHK(flush_timing(ctx));
    CHK(read_status(ctx));
    ctx->state_set = true;
    return QTP_OK;
}

int qtp_get_state(qtp_ctx *ctx, double *q, double *p, double *F, double *fphys, double *disp)
{
    if (!ctx) return QTP_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int P = ctx->P;
    if (q) CHK(download_planar(ctx, ctx->q, P, q));
    if (p) CHK(download_planar(ctx, ctx->p, P, p));
    if (F) CHK(download_planar(ctx, ctx->F, P, F));
    if (fphys) CHK(download_planar(ctx, ctx->fph, P, fphys));
    if (disp) CHK(download_disp(ctx, disp, 0, P));
    return QTP_OK;
}

// ---- checkpoint / restart ------------------------------------------------------

// File layout (little endian): "QTPCKPT1", int32 version (2), int32 n_mols,
// n_beads, list_size, k_max; double boxl[3], T_K, dtfs, eps_ewald,
// epsilon_sc, kappa, r_c, k_cut (the effective Ewald parameters), the LJ
// cutoff; uint64
// seed, step counter; double heat; then in the reference layout: q, p
// (n_beads * n_dof), disp (n_beads * 3 * n_mols), point (n_beads * n_mols
// int32), list (n_beads * list_size int32), rebuild counts (n_beads doubles).
// A load into a context whose system, temperature, time step or force
// parameters differ is rejected: the run would not continue as saved.
static const char QTP_CKPT_MAGIC[8] = {'Q', 'T', 'P', 'C', 'K', 'P', 'T', '1'};
static const int32_t QTP_CKPT_VERSION = 2;
static const int QTP_CKPT_NPAR = 11;

static void ckpt_header(const qtp_ctx *ctx, int32_t hdr[5], double par[QTP_CKPT_NPAR])
{
    hdr[0] = QTP_CKPT_VERSION;
    hdr[1] = ctx->n_mols;
    hdr[2] = ctx->P;
    hdr[3] = ctx->list_size;
    hdr[4] = ctx->drv.k_max;
    const double v[QTP_CKPT_NPAR] = {ctx->drv.boxl[0], ctx->drv.boxl[1], ctx->drv.boxl[2],
                                     ctx->prm.T_K,     ctx->prm.dtfs,    ctx->prm.eps_ewald,
                                     ctx->prm.epsilon_sc, ctx->drv.kappa, ctx->drv.rcut,
                                     ctx->drv.k_cut, ctx->drv.lj_rcut};
    for (int i = 0; i < QTP_CKPT_NPAR; ++i) par[i] = v[i];
}

int qtp_save_checkpoint(qtp_ctx *ctx, const char *path)
{
    if (!ctx || !path) return QTP_ERR_ARG;
    if (!ctx->state_set) {
        ctx->err = "qtp_save_checkpoint: no resident state";
        return QTP_ERR_STATE;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int P = ctx->P;
    const size_t nd = (size_t)P * ctx->n_dof, nm = (size_t)P * 3 * ctx->n_mols;
    std::vector<double> q(nd), p(nd), disp(nm), nreb(P);
    std::vector<int> point((size_t)P * ctx->n_mols), list((size_t)P * ctx->list_size);
    CHK(qtp_get_state(ctx, q.data(), p.data(), nullptr, nullptr, disp.data()));
    CHK(qtp_get_lists(ctx, point.data(), list.data()));
    unsigned long long ctr = 0;
    double heat = 0.0;
    HIPCHK(hipMemcpy(&ctr, ctx->step_ctr, sizeof(ctr), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&heat, ctx->heat, sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(nreb.data(), ctx->nreb, P * sizeof(double), hipMemcpyDeviceToHost));
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        ctx->err = std::string("cannot open ") + path;
        return QTP_ERR_ARG;
    }
    int32_t hdr[5];
    double par[QTP_CKPT_NPAR];
    ckpt_header(ctx, hdr, par);
    const uint64_t u[2] = {ctx->seed, ctr};
    bool ok = std::fwrite(QTP_CKPT_MAGIC, 1, 8, f) == 8 && std::fwrite(hdr, 4, 5, f) == 5 &&
              std::fwrite(par, 8, QTP_CKPT_NPAR, f) == (size_t)QTP_CKPT_NPAR &&
              std::fwrite(u, 8, 2, f) == 2 &&
              std::fwrite(&heat, 8, 1, f) == 1 && std::fwrite(q.data(), 8, nd, f) == nd &&
              std::fwrite(p.data(), 8, nd, f) == nd && std::fwrite(disp.data(), 8, nm, f) == nm &&
              std::fwrite(point.data(), 4, point.size(), f) == point.size() &&
              std::fwrite(list.data(), 4, list.size(), f) == list.size() &&
              std::fwrite(nreb.data(), 8, P, f) == (size_t)P;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        ctx->err = std::string("write failed: ") + path;
        return QTP_ERR_ARG;
    }
    return QTP_OK;
}

int qtp_load_checkpoint(qtp_ctx *ctx, const char *path, const double *charge, const double *mass)
{
    if (!ctx || !path) return QTP_ERR_ARG;
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        ctx->err = std::string("cannot open ") + path;
        return QTP_ERR_ARG;
    }
    const int P = ctx->P;
    const size_t nd = (size_t)P * ctx->n_dof, nm = (size_t)P * 3 * ctx->n_mols;
    char magic[8];
    int32_t hdr[5], want_hdr[5];
    double par[QTP_CKPT_NPAR], want_par[QTP_CKPT_NPAR], heat = 0.0;
    ckpt_header(ctx, want_hdr, want_par);
    uint64_t u[2];
    std::vector<double> q(nd), p(nd), disp(nm), nreb(P);
    std::vector<int> point((size_t)P * ctx->n_mols), list((size_t)P * ctx->list_size);
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, QTP_CKPT_MAGIC, 8) == 0 &&
              std::fread(hdr, 4, 5, f) == 5 && hdr[0] == QTP_CKPT_VERSION &&
              hdr[1] == ctx->n_mols && hdr[2] == P && hdr[3] == ctx->list_size &&
              std::fread(par, 8, QTP_CKPT_NPAR, f) == (size_t)QTP_CKPT_NPAR &&
              std::fread(u, 8, 2, f) == 2 && std::fread(&heat, 8, 1, f) == 1 &&
              std::fread(q.data(), 8, nd, f) == nd && std::fread(p.data(), 8, nd, f) == nd &&
              std::fread(disp.data(), 8, nm, f) == nm &&
              std::fread(point.data(), 4, point.size(), f) == point.size() &&
              std::fread(list.data(), 4, list.size(), f) == list.size() &&
              std::fread(nreb.data(), 8, P, f) == (size_t)P;
    std::fclose(f);
    if (!ok) {
        ctx->err = std::string("not a checkpoint of this system: ") + path;
        return QTP_ERR_ARG;
    }
    static const char *const what[QTP_CKPT_NPAR] = {
        "box", "box", "box", "T_K", "dtfs", "eps_ewald", "epsilon_sc", "Ewald kappa",
        "Ewald r_c", "Ewald k_cut", "LJ cutoff"};
    for (int k = 0; k < QTP_CKPT_NPAR; ++k)
        if (par[k] != want_par[k]) {
            ctx->err = std::string("checkpoint ") + what[k] + " differs from the context's";
            return QTP_ERR_ARG;
        }
    if (hdr[4] != want_hdr[4]) {
        ctx->err = "checkpoint Ewald k_max differs from the context's";
        return QTP_ERR_ARG;
    }
    SplineScope sp(ctx, ctx->spl_resident);
    CHK(check_finite(ctx, q.data(), (long)nd, "positions"));
    CHK(check_finite(ctx, p.data(), (long)nd, "momenta"));
    invalidate_graphs(ctx);
    CHK(upload_charges_masses(ctx, charge, mass));
    CHK(upload_planar(ctx, ctx->q, q.data(), P));
    CHK(upload_planar(ctx, ctx->p, p.data(), P));
    CHK(upload_disp(ctx, disp.data(), 0, P));
    // the saved half lists, and the full lists expanded from them: each row in
    // ascending partner order, as the list kernels write it, so the force
    // sums run in the same order as before the stop
    CHK(upload_lists(ctx, point.data(), list.data(), 0, P, 0));
    HIPCHK(hipMemset(ctx->status, 0, sizeof(int)));
    if (P >= 2) {
        CHK(run_sc_forces(ctx, 0, nullptr));
    } else {
        CHK(class_initial_forces(ctx));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    CHK(flush_timing(ctx));
    CHK(read_status(ctx));
    HIPCHK(hipMemcpy(ctx->heat, &heat, sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->nreb, nreb.data(), P * sizeof(double), hipMemcpyHostToDevice));
    const unsigned long long ctr = u[1];
    HIPCHK(hipMemcpy(ctx->step_ctr, &ctr, sizeof(ctr), hipMemcpyHostToDevice));
    ctx->seed = u[0];
    ctx->state_set = true;
    return QTP_OK;
}

int qtp_get_lists(qtp_ctx *ctx, int *point, int *list)
{
    if (!ctx || !point || !list) return QTP_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(point, ctx->point, (size_t)ctx->P * ctx->n_mols * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(list, ctx->list, (size_t)ctx->P * ctx->list_size * sizeof(int), hipMemcpyDeviceToHost));
    return QTP_OK;
}

int qtp_set_seed(qtp_ctx *ctx, uint64_t seed)
{
    if (!ctx) return QTP_ERR_ARG;
    ctx->seed = seed;
    invalidate_graphs(ctx);
    return QTP_OK;
}

int qtp_get_virial(qtp_ctx *ctx, double *virial, double *energy)
{
    if (!ctx || !virial) return QTP_ERR_ARG;
    if (!ctx->state_set) {
        ctx->err = "qtp_get_virial: no resident state (call qtp_set_state)";
        return QTP_ERR_STATE;
    }
    SplineScope sp(ctx, ctx->spl_resident);
    const int P = ctx->P;
    // physical forces go to scratch: the resident force arrays are left as they are
    CHK(run_phase(ctx, ctx->q, P, 0, 0, ctx->fpm, ctx->Ve, ctx->vout));
    HIPCHK(hipMemcpyAsync(virial, ctx->vout, (size_t)P * 6 * sizeof(double),
                          hipMemcpyDeviceToHost, ctx->stream));
    if (energy)
        HIPCHK(hipMemcpyAsync(energy, ctx->Ve, (size_t)P * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    CHK(flush_timing(ctx));
    return read_status(ctx);
}

static int step_once(qtp_ctx *ctx, int nvt)
{
    SplineScope sp(ctx, ctx->spl_resident);
    if (ctx->P >= 2) return run_sc_step(ctx, nvt, 0, ctx->step_ctr);
    return run_class_step(ctx, ctx->step_ctr, true, nvt);
}

// NVT flag of a resident step: SC steps always take it; classical steps only
// with the Langevin thermostat (PILE-L kind), the one they implement.
static int resident_nvt(qtp_ctx *ctx, int is_nvt, int *nvt)
{
    *nvt = is_nvt ? 1 : 0;
    if (*nvt && ctx->P == 1 && !(ctx->cthermo == QTP_THERMO_PILE_L && ctx->tau_fs > 0.0)) {
        ctx->err = "classical NVT steps take the Langevin thermostat: "
                   "qtp_set_thermostat(ctx, QTP_THERMO_PILE_L, tau_fs > 0)";
        return QTP_ERR_STATE;
    }
    return QTP_OK;
}

// One step of the resident state captured as a graph (not instantiated).
static int capture_step(qtp_ctx *ctx, int nvt, hipGraph_t *out)
{
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int st = step_once(ctx, nvt);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (st != QTP_OK) {
        if (g) hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) {
        ctx->err = std::string("graph capture: ") + hipGetErrorString(e);
        return QTP_ERR_HIP;
    }
    *out = g;
    return QTP_OK;
}

int qtp_run(qtp_ctx *ctx, int n_steps, int is_nvt)
{
    if (!ctx || n_steps < 0) return QTP_ERR_ARG;
    if (!ctx->state_set) {
        ctx->err = "qtp_run: no resident state (call qtp_set_state)";
        return QTP_ERR_STATE;
    }
    int nvt = 0;
    CHK(resident_nvt(ctx, is_nvt, &nvt));
    if (ctx->timing) {
        for (int i = 0; i < n_steps; ++i) CHK(step_once(ctx, nvt));
        CHK(flush_timing(ctx));
        return QTP_OK;
    }
    if (!ctx->graph[nvt]) {
        hipGraph_t g;
        CHK(capture_step(ctx, nvt, &g));
        hipError_t ei = hipGraphInstantiate(&ctx->graph[nvt], g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ei != hipSuccess) {
            ctx->err = std::string("graph instantiate: ") + hipGetErrorString(ei);
            ctx->graph[nvt] = nullptr;
            return QTP_ERR_HIP;
        }
    }
    for (int i = 0; i < n_steps; ++i) HIPCHK(hipGraphLaunch(ctx->graph[nvt], ctx->stream));
    return QTP_OK;
}

// ---- batched replicas ----------------------------------------------------------

struct qtp_batch {
    std::vector<qtp_ctx *> ctx;
    std::vector<int> nvt;
    std::string err;
};

// Each replica keeps its own captured step graph and its own streams; a batch
// step launches every replica's graph on its replica's stream, so the GPU
// overlaps the replicas (the runtime executes the branches of one combined
// graph, flat or of child graphs, one after another: measured 1.0x, against
// 1.5x for per-stream launches at 8 x (216 H2O x 32 beads)).
int qtp_batch_create(qtp_ctx *const *ctxs, const int *is_nvt, int n, qtp_batch **out)
{
    if (!ctxs || n < 1 || !out) return QTP_ERR_ARG;
    qtp_batch *b = new qtp_batch();
    *out = b;
    for (int i = 0; i < n; ++i) {
        qtp_ctx *c = ctxs[i];
        if (!c) {
            b->err = "null context in batch";
            return QTP_ERR_ARG;
        }
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == c) {
                b->err = "context listed twice in batch";
                return QTP_ERR_ARG;
            }
        if (!c->state_set || c->timing) {
            b->err = "batch member without resident state (or timing on)";
            return QTP_ERR_STATE;
        }
        int nvt = 0;
        if (resident_nvt(c, is_nvt && is_nvt[i], &nvt) != QTP_OK) {
            b->err = "replica " + std::to_string(i) + ": " + c->err;
            return QTP_ERR_STATE;
        }
        // capture and instantiate the replica's step graph now (zero steps run)
        const int st = qtp_run(c, 0, nvt);
        if (st != QTP_OK) {
            b->err = "replica " + std::to_string(i) + ": " + c->err;
            return st;
        }
        b->ctx.push_back(c);
        b->nvt.push_back(nvt);
    }
    return QTP_OK;
}

int qtp_batch_run(qtp_batch *b, int n_steps)
{
    if (!b || b->ctx.empty() || n_steps < 0) return QTP_ERR_ARG;
    for (int s = 0; s < n_steps; ++s)
        for (size_t i = 0; i < b->ctx.size(); ++i) {
            qtp_ctx *c = b->ctx[i];
            if (!c->graph[b->nvt[i]] ||
                hipGraphLaunch(c->graph[b->nvt[i]], c->stream) != hipSuccess) {
                b->err = "replica " + std::to_string(i) + ": graph launch failed";
                return QTP_ERR_HIP;
            }
        }
    return QTP_OK;
}

int qtp_batch_sync(qtp_batch *b)
{
    if (!b) return QTP_ERR_ARG;
    for (size_t i = 0; i < b->ctx.size(); ++i) {
        const int st = qtp_sync(b->ctx[i]);
        if (st != QTP_OK) {
            b->err = "replica " + std::to_string(i) + ": " + b->ctx[i]->err;
            return st;
        }
    }
    return QTP_OK;
}

const char *qtp_batch_last_error(const qtp_batch *b) { return b ? b->err.c_str() : "null batch"; }

void qtp_batch_destroy(qtp_batch *b)
{
    if (!b) return;
    for (qtp_ctx *c : b->ctx) hipStreamSynchronize(c->stream);
    delete b;
}

// ---- resident observables ---------------------------------------------------

// Observables of the resident state into dst (every == 0) or, on steps whose
// counter is a multiple of every, into the next trace row.
static int launch_obs(qtp_ctx *ctx, int every, double *dst)
{
    const int nb = (3 * ctx->NAP + 255) / 256;
    k_obs_part<<<nb, 256, 0, ctx->stream>>>(ctx->P, ctx->n_atoms, ctx->NAP, ctx->drv.omega_psq,
                                             ctx->q, ctx->p, ctx->fph, ctx->m_atom, ctx->obs);
    k_obs_fin<<<1, 64, 0, ctx->stream>>>(ctx->P, nb, 0.5 * ctx->n_dof / ctx->drv.beta, ctx->obs,
                                          ctx->out, ctx->step_ctr, every, ctx->trace_cap,
                                          ctx->ntrace, dst);
    HIPCHK(hipGetLastError());
    return QTP_OK;
}

int qtp_get_observables(qtp_ctx *ctx, double out[4])
{
    if (!ctx || !out) return QTP_ERR_ARG;
    if (!ctx->state_set) {
        ctx->err = "qtp_get_observables: no resident state (call qtp_set_state)";
        return QTP_ERR_STATE;
    }
    CHK(launch_obs(ctx, 0, ctx->obs_out));
    HIPCHK(hipMemcpyAsync(out, ctx->obs_out, 4 * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return read_status(ctx);
}

int qtp_run_observed(qtp_ctx *ctx, int n_steps, int is_nvt, int every, double *trace,
                     int *n_samples)
{
    if (!ctx || n_steps < 0 || every < 1 || !trace) return QTP_ERR_ARG;
    if (!ctx->state_set) {
        ctx->err = "qtp_run_observed: no resident state (call qtp_set_state)";
        return QTP_ERR_STATE;
    }
    int nvt = 0;
    CHK(resident_nvt(ctx, is_nvt, &nvt));
    const int need = n_steps / every + 1;
    if (need > ctx->trace_cap) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        invalidate_graphs(ctx);
        if (ctx->trace) hipFree(ctx->trace);
        ctx->trace = nullptr;
        ctx->trace_cap = 0;
        CHK(dalloc(ctx, &ctx->trace, (size_t)4 * need));
        ctx->trace_cap = need;
    }
    if (every != ctx->obs_every) {
        for (auto &g : ctx->graph_obs)
            if (g) {
                hipGraphExecDestroy(g);
                g = nullptr;
            }
        ctx->obs_every = every;
    }
    HIPCHK(hipMemsetAsync(ctx->ntrace, 0, sizeof(int), ctx->stream));
    if (ctx->timing) {
        for (int i = 0; i < n_steps; ++i) {
            CHK(step_once(ctx, nvt));
            CHK(launch_obs(ctx, every, ctx->trace));
        }
        CHK(flush_timing(ctx));
    } else {
        if (!ctx->graph_obs[nvt]) {
            HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            int st = step_once(ctx, nvt);
            if (st == QTP_OK) st = launch_obs(ctx, every, ctx->trace);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
            if (st != QTP_OK) {
                if (g) hipGraphDestroy(g);
                return st;
            }
            if (e != hipSuccess) {
                ctx->err = std::string("graph capture: ") + hipGetErrorString(e);
                return QTP_ERR_HIP;
            }
            const hipError_t ei = hipGraphInstantiate(&ctx->graph_obs[nvt], g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (ei != hipSuccess) {
                ctx->graph_obs[nvt] = nullptr;
                ctx->err = std::string("graph instantiate: ") + hipGetErrorString(ei);
                return QTP_ERR_HIP;
            }
        }
        for (int i = 0; i < n_steps; ++i) HIPCHK(hipGraphLaunch(ctx->graph_obs[nvt], ctx->stream));
    }
    int n = 0;
    HIPCHK(hipMemcpyAsync(&n, ctx->ntrace, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    n = std::min(n, need);
    if (n > 0)
        HIPCHK(hipMemcpy(trace, ctx->trace, (size_t)4 * n * sizeof(double), hipMemcpyDeviceToHost));
    if (n_samples) *n_samples = n;
    return read_status(ctx);
}

int qtp_sync(qtp_ctx *ctx)
{
    if (!ctx) return QTP_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return read_status(ctx);
}

int qtp_get_scalars(qtp_ctx *ctx, double out[6])
{
    if (!ctx || !out) return QTP_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double o[2], h;
    unsigned long long n;
    int st;
    std::vector<double> nr(ctx->P);
    HIPCHK(hipMemcpy(o, ctx->out, 2 * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&h, ctx->heat, sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&n, ctx->step_ctr, sizeof(n), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&st, ctx->status, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(nr.data(), ctx->nreb, ctx->P * sizeof(double), hipMemcpyDeviceToHost));
    double tot = 0.0;
    for (double v : nr) tot += v;
    out[0] = o[0];
    out[1] = o[1];
    out[2] = h;
    out[3] = tot;
    out[4] = (double)n;
    out[5] = (double)st;
    return QTP_OK;
}

int qtp_enable_timing(qtp_ctx *ctx, int on)
{
    if (!ctx) return QTP_ERR_ARG;
    ctx->timing = on != 0;
    ctx->timers.clear();
    return QTP_OK;
}

int qtp_get_timing(qtp_ctx *ctx, const char *kernel, double *ms_total, int *launches)
{
    if (!ctx || !kernel) return QTP_ERR_ARG;
    CHK(flush_timing(ctx));
    auto it = ctx->timers.find(kernel);
    if (ms_total) *ms_total = it == ctx->timers.end() ? 0.0 : it->second.ms;
    if (launches) *launches = it == ctx->timers.end() ? 0 : it->second.n;
    return QTP_OK;
}

// ---- learned centroid-force correction --------------------------------------------

int qtp_nn_descriptor_params(double *out, int n)
{
    // r_rad, r_ang, mu0, dmu, eta_r (Angstrom), zeta[16], lambda[16], eta_a[16],
    // w[2][2] (element-pair weights, O = 0, H = 1)
    std::vector<double> v = {RR, RA, MU0, DMU, ETAR};
    for (int t = 0; t < NANG; ++t) v.push_back((double)(1 << ((t >> 1) & 3)));
    for (int t = 0; t < NANG; ++t) v.push_back((t & 1) ? -1.0 : 1.0);
    for (int t = 0; t < NANG; ++t) v.push_back((t >> 3) ? (double)ETA_A1 : (double)ETA_A0);
    v.push_back(1.0);
    v.push_back(0.5);
    v.push_back(1.0);
    v.push_back(0.5);
    if (out)
        for (int t = 0; t < std::min(n, (int)v.size()); ++t) out[t] = v[t];
    return (int)v.size();
}

int qtp_nn_create(int n_mols, const double boxl[3], int max_frames, qtp_nn **out)
{
    if (!out || n_mols < 2 || !boxl || max_frames < 1) return QTP_ERR_ARG;
    qtp_nn *nn = new qtp_nn();
    *out = nn;
    nn->n_mols = n_mols;
    nn->n_atoms = 3 * n_mols;
    nn->NAP = ((nn->n_atoms + 63) / 64) * 64;
    nn->max_frames = max_frames;
    for (int k = 0; k < 3; ++k) {
        if (!(boxl[k] > 0.0) || 2.0 * RR / ANGD > boxl[k]) {
            nn->err = "box edges must exceed twice the 6 A descriptor cutoff";
            return QTP_ERR_ARG;
        }
        nn->box.L[k] = boxl[k];
        nn->box.inv[k] = 1.0 / boxl[k];
    }
    NNCHK(hipStreamCreate(&nn->stream));
    const long NA = (long)max_frames * nn->n_atoms, N3 = 3L * nn->NAP * max_frames;
    const NNBatch b = nn_batch(nn, max_frames);
    int st = QTP_OK;
    if ((st = nalloc(nn, &nn->theta, 2 * NPE)) || (st = nalloc(nn, &nn->am, 2 * NPE)) ||
        (st = nalloc(nn, &nn->av, 2 * NPE)) || (st = nalloc(nn, &nn->grad, 2 * NPE)) ||
        (st = nalloc(nn, &nn->w1, 2 * W1E)) || (st = nalloc(nn, &nn->w1t, 2 * W1E)) ||
        (st = nalloc(nn, &nn->w2, 2 * W2E)) || (st = nalloc(nn, &nn->w2t, 2 * W2E)) ||
        (st = nalloc(nn, &nn->b1, 2 * NH)) || (st = nalloc(nn, &nn->b2, 2 * NH)) ||
        (st = nalloc(nn, &nn->w3, 2 * NH)) || (st = nalloc(nn, &nn->b3, 2)) ||
        (st = nalloc(nn, &nn->mean, 2 * NIN)) || (st = nalloc(nn, &nn->stdv, 2 * NIN)) ||
        (st = nalloc(nn, &nn->pos, N3)) || (st = nalloc(nn, &nn->tgt, N3)) ||
        (st = nalloc(nn, &nn->fwork, N3)) || (st = nalloc(nn, &nn->dsum, max_frames)) ||
        (st = nalloc(nn, &nn->gt, NA * NIN)) || (st = nalloc(nn, &nn->ut, NA * NIN)) ||
        (st = nalloc(nn, &nn->dEdG, NA * NIN)) || (st = nalloc(nn, &nn->graw, NA * NIN)) ||
        (st = nalloc(nn, &nn->Eatom, NA)) || (st = nalloc(nn, &nn->res, 3 * NA)) ||
        (st = nalloc(nn, &nn->sq, NA)) || (st = nalloc(nn, &nn->nbr, NA * NMAX)) ||
        (st = nalloc(nn, &nn->nn, NA)) || (st = nalloc(nn, &nn->nba, NA * NAMAX)) ||
        (st = nalloc(nn, &nn->nna, NA)) || (st = nalloc(nn, &nn->status, 1)) ||
        (st = nalloc(nn, &nn->dang, NA * NAMAX)) || (st = nalloc(nn, &nn->dself, NA)) ||
        (st = nalloc(nn, &nn->P4, NA)) || (st = nalloc(nn, &nn->Etile, b.to + b.th)) ||
        (st = nalloc(nn, &nn->tp, (size_t)(b.tiles_o + b.tiles_h) * 3 * NH)) ||
        (st = nalloc(nn, &nn->part, (size_t)(b.nco + b.nch) * WG_PART)) ||
        (st = nalloc(nn, &nn->inter, nn_inter_elems(b))))
        return st;
    std::vector<float> one(2 * NIN, 1.0f);
    NNCHK(hipMemcpy(nn->stdv, one.data(), one.size() * sizeof(float), hipMemcpyHostToDevice));
    NNCHK(hipFuncSetAttribute((const void *)k_nn_train_fwd,
                              hipFuncAttributeMaxDynamicSharedMemorySize, TRAIN_LDS));
    return QTP_OK;
}

void qtp_nn_destroy(qtp_nn *nn)
{
    if (!nn) return;
    if (nn->stream) hipStreamSynchronize(nn->stream);
    void *ptrs[] = {nn->theta, nn->am, nn->av, nn->grad, nn->w1, nn->w1t, nn->w2, nn->w2t,
                    nn->b1, nn->b2, nn->w3, nn->b3, nn->mean, nn->stdv, nn->pos, nn->tgt,
                    nn->fwork, nn->dsum, nn->gt, nn->ut, nn->dEdG, nn->graw, nn->Eatom,
                    nn->res, nn->sq, nn->nbr, nn->nn, nn->nba, nn->nna, nn->status, nn->tp,
                    nn->dang, nn->dself, nn->P4, nn->Etile,
                    nn->part, nn->inter, nn->dpos, nn->dtgt};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (nn->stream) hipStreamDestroy(nn->stream);
    delete nn;
}

const char *qtp_nn_last_error(const qtp_nn *nn) { return nn ? nn->err.c_str() : "null model"; }

int qtp_nn_set_model(qtp_nn *nn, const float *theta, const float *mean, const float *stdv,
                     double e_scale)
{
    if (!nn || !theta || !mean || !stdv || !(e_scale > 0.0)) return QTP_ERR_ARG;
    for (int t = 0; t < 2 * NPE; ++t)
        if (!std::isfinite(theta[t])) {
            nn->err = "parameters contain a non-finite value";
            return QTP_ERR_NONFINITE;
        }
    for (int t = 0; t < 2 * NIN; ++t)
        if (!std::isfinite(mean[t]) || !(stdv[t] > 0.0f)) {
            nn->err = "normalisation: finite means and positive deviations required";
            return QTP_ERR_ARG;
        }
    NNCHK(hipStreamSynchronize(nn->stream));
    NNCHK(hipMemcpy(nn->theta, theta, 2 * NPE * sizeof(float), hipMemcpyHostToDevice));
    NNCHK(hipMemcpy(nn->mean, mean, 2 * NIN * sizeof(float), hipMemcpyHostToDevice));
    NNCHK(hipMemcpy(nn->stdv, stdv, 2 * NIN * sizeof(float), hipMemcpyHostToDevice));
    NNCHK(hipMemset(nn->am, 0, 2 * NPE * sizeof(float)));
    NNCHK(hipMemset(nn->av, 0, 2 * NPE * sizeof(float)));
    nn->adam_t = 0;
    nn->e_scale = e_scale;
    k_nn_update<<<(2 * NPE + 255) / 256, 256, 0, nn->stream>>>(2, nullptr, 0, 0, nullptr, 0, 0, 0.0,
                                                               nn->theta, nn->am, nn->av,
                                                               AdamArgs{}, nullptr, nn_dev(nn));
    NNCHK(hipGetLastError());
    NNCHK(hipStreamSynchronize(nn->stream));
    nn->model_set = true;
    return QTP_OK;
}

int qtp_nn_get_model(qtp_nn *nn, float *theta)
{
    if (!nn || !theta) return QTP_ERR_ARG;
    NNCHK(hipStreamSynchronize(nn->stream));
    NNCHK(hipMemcpy(theta, nn->theta, 2 * NPE * sizeof(float), hipMemcpyDeviceToHost));
    return QTP_OK;
}

static int nn_need_model(qtp_nn *nn)
{
    if (!nn->model_set) {
        nn->err = "no model (qtp_nn_set_model)";
        return QTP_ERR_STATE;
    }
    return QTP_OK;
}

int qtp_nn_eval(qtp_nn *nn, int n_frames, const double *pos, double *energy, double *forces,
                float *descriptors)
{
    if (!nn || n_frames < 1 || !pos) return QTP_ERR_ARG;
    CHK(nn_need_model(nn));
    const long nd = 3L * nn->n_atoms, N3 = 3L * nn->NAP;
    for (int f0 = 0; f0 < n_frames; f0 += nn->max_frames) {
        const int F = std::min(nn->max_frames, n_frames - f0);
        for (long t = 0; t < F * nd; ++t)
            if (!std::isfinite(pos[f0 * nd + t])) {
                nn->err = "positions contain a non-finite value";
                return QTP_ERR_NONFINITE;
            }
        CHK(nn_upload_frames(nn, nn->pos, pos + f0 * nd, F));
        const NNBatch b = nn_batch(nn, F);
        const NNGeom g = nn_geom(nn, F);
        hipStream_t s = nn->stream;
        NNCHK(hipMemsetAsync(nn->status, 0, sizeof(int), s));
        NNCHK(hipMemsetAsync(nn->fwork, 0, (size_t)F * N3 * sizeof(double), s));
        CHK(nn_forward(nn, s, nn->pos, F, b, nn->status, descriptors ? nn->graw : nullptr));
        k_nn_force<0><<<b.nblk_atom, AT, 0, s>>>(g, nn->P4, nn->nbr, nn->nn, nn->nba, nn->nna,
                                                 nn->dEdG, nn->dang, nn->dself, nn->fwork, nullptr,
                                                 nullptr, nullptr);
        for (int f = 0; f < F; ++f)
            k_nn_sum<<<1, 1024, 0, s>>>(nn->n_atoms, nn->Eatom + (long)f * nn->n_atoms, 1.0, 0,
                                        nn->dsum + f);
        NNCHK(hipGetLastError());
        NNCHK(hipStreamSynchronize(s));
        CHK(nn_read_status(nn));
        if (energy) NNCHK(hipMemcpy(energy + f0, nn->dsum, F * sizeof(double), hipMemcpyDeviceToHost));
        if (forces) {
            std::vector<double> h((size_t)F * N3);
            NNCHK(hipMemcpy(h.data(), nn->fwork, h.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (int f = 0; f < F; ++f)
                for (int a = 0; a < nn->n_atoms; ++a)
                    for (int d = 0; d < 3; ++d)
                        forces[((long)(f0 + f) * nn->n_atoms + a) * 3 + d] =
                            h[f * N3 + (long)d * nn->NAP + a];
        }
        if (descriptors)
            NNCHK(hipMemcpy(descriptors + (long)f0 * nn->n_atoms * NIN, nn->graw,
                            (size_t)F * nn->n_atoms * NIN * sizeof(float), hipMemcpyDeviceToHost));
    }
    return QTP_OK;
}

int qtp_nn_set_data(qtp_nn *nn, int n_frames, const double *pos, const double *target)
{
    if (!nn || n_frames < 1 || !pos || !target) return QTP_ERR_ARG;
    const long nd = 3L * nn->n_atoms, N3 = 3L * nn->NAP;
    for (long t = 0; t < n_frames * nd; ++t)
        if (!std::isfinite(pos[t]) || !std::isfinite(target[t])) {
            nn->err = "training data contain a non-finite value";
            return QTP_ERR_NONFINITE;
        }
    NNCHK(hipStreamSynchronize(nn->stream));
    if (nn->dpos) hipFree(nn->dpos);
    if (nn->dtgt) hipFree(nn->dtgt);
    nn->dpos = nn->dtgt = nullptr;
    nn->n_data = 0;
    CHK(nalloc(nn, &nn->dpos, (size_t)n_frames * N3));
    CHK(nalloc(nn, &nn->dtgt, (size_t)n_frames * N3));
    CHK(nn_upload_frames(nn, nn->dpos, pos, n_frames));
    CHK(nn_upload_frames(nn, nn->dtgt, target, n_frames));
    nn->n_data = n_frames;
    return QTP_OK;
}

int qtp_nn_loss_grad(qtp_nn *nn, int frame0, int n_frames, double *loss, float *grad)
{
    if (!nn || frame0 < 0 || n_frames < 1 || n_frames > nn->max_frames) return QTP_ERR_ARG;
    CHK(nn_need_model(nn));
    if (frame0 + n_frames > nn->n_data) {
        nn->err = "frames beyond the training data (qtp_nn_set_data)";
        return QTP_ERR_ARG;
    }
    hipStream_t s = nn->stream;
    NNCHK(hipMemsetAsync(nn->status, 0, sizeof(int), s));
    CHK(nn_train_batch(nn, s, frame0, n_frames, 0, 0.0, nn->dsum));
    NNCHK(hipStreamSynchronize(s));
    CHK(nn_read_status(nn));
    if (loss) NNCHK(hipMemcpy(loss, nn->dsum, sizeof(double), hipMemcpyDeviceToHost));
    if (grad) NNCHK(hipMemcpy(grad, nn->grad, 2 * NPE * sizeof(float), hipMemcpyDeviceToHost));
    return QTP_OK;
}

int qtp_nn_train(qtp_nn *nn, int n_steps, int batch_frames, double lr, uint64_t seed,
                 double *loss_trace)
{
    if (!nn || n_steps < 0 || batch_frames < 1 || batch_frames > nn->max_frames || !(lr > 0.0))
        return QTP_ERR_ARG;
    CHK(nn_need_model(nn));
    if (nn->n_data < batch_frames) {
        nn->err = "fewer training frames than one batch (qtp_nn_set_data)";
        return QTP_ERR_ARG;
    }
    const int nb = nn->n_data / batch_frames;
    double *trace = nullptr;
    CHK(nalloc(nn, &trace, std::max(1, n_steps)));
    std::mt19937_64 rng(seed);
    std::vector<int> order(nb);
    hipStream_t s = nn->stream;
    NNCHK(hipMemsetAsync(nn->status, 0, sizeof(int), s));
    int st = QTP_OK;
    for (int t = 0; t < n_steps && st == QTP_OK; ++t) {
        if (t % nb == 0) {
            for (int q = 0; q < nb; ++q) order[q] = q;
            std::shuffle(order.begin(), order.end(), rng);
        }
        st = nn_train_batch(nn, s, order[t % nb] * batch_frames, batch_frames, 1, lr, trace + t);
    }
    if (st == QTP_OK) {
        if (hipStreamSynchronize(s) != hipSuccess) st = QTP_ERR_HIP;
        else st = nn_read_status(nn);
    }
    if (st == QTP_OK && loss_trace && n_steps > 0 &&
        hipMemcpy(loss_trace, trace, n_steps * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        st = QTP_ERR_HIP;
    hipFree(trace);
    return st;
}

int qtp_set_nn(qtp_ctx *ctx, qtp_nn *nn)
{
    if (!ctx) return QTP_ERR_ARG;
    if (nn) {
        if (ctx->P != 1) {
            ctx->err = "the learned correction drives classical (centroid) dynamics: n_beads = 1";
            return QTP_ERR_ARG;
        }
        if (nn->n_mols != ctx->n_mols || nn->box.L[0] != ctx->box.L[0] ||
            nn->box.L[1] != ctx->box.L[1] || nn->box.L[2] != ctx->box.L[2]) {
            ctx->err = "model made for another system (molecule count or box)";
            return QTP_ERR_ARG;
        }
        if (!nn->model_set) {
            ctx->err = "model without parameters (qtp_nn_set_model)";
            return QTP_ERR_STATE;
        }
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (nn && !ctx->fnn) CHK(dalloc(ctx, &ctx->fnn, 3 * (size_t)ctx->NAP));
    invalidate_graphs(ctx);
    ctx->nn = nn;
    ctx->state_set = false;
    return QTP_OK;
}

// ---- launch configuration and autotuner -----------------------------------------------

int qtp_get_tuning(const qtp_ctx *ctx, int cfg[9])
{
    if (!ctx || !cfg) return QTP_ERR_ARG;
    const TuneCfg c = current_tuning(ctx);
    cfg[0] = c.lpm;
    cfg[1] = c.kmf;
    cfg[2] = c.carg;
    cfg[3] = c.small;
    cfg[4] = c.usplit;
    cfg[5] = c.fused;
    cfg[6] = c.rsw;
    cfg[7] = c.rsplit;
    cfg[8] = c.cpb;
    return QTP_OK;
}

int qtp_set_tuning(qtp_ctx *ctx, const int cfg[9])
{
    if (!ctx || !cfg) return QTP_ERR_ARG;
    TuneCfg c;
    c.lpm = cfg[0];
    c.kmf = cfg[1];
    c.carg = cfg[2];
    c.small = cfg[3];
    c.usplit = cfg[4];
    c.fused = cfg[5];
    c.rsw = cfg[6];
    c.rsplit = cfg[7];
    c.cpb = cfg[8];
    if (!tuning_valid(ctx, c)) {
        ctx->err = "tuning: lanes 16 or 64, matrix k-space 0/1 (1 only for k_max <= 23), "
                   "registers and 16-wave blocks -1/0/1, blocks per chunk 0-4, fused "
                   "propagation -1/0/1, real-space waves 2 or 4, real-space eighths on the "
                   "second stream 0-4, structure-factor chunks per block 0/1/2/4/8/16";
        return QTP_ERR_ARG;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    apply_tuning(ctx, c);
    SplineScope sp(ctx, ctx->spl_resident);
    return tune_refresh(ctx);
}

int qtp_autotune(qtp_ctx *ctx, int flags, double report[11])
{
    if (!ctx) return QTP_ERR_ARG;
    if (!ctx->state_set) {
        ctx->err = "autotune: set the state first (qtp_set_state)";
        return QTP_ERR_STATE;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    SplineScope sp(ctx, ctx->spl_resident);
    const bool timing = ctx->timing;
    ctx->timing = false;
    const std::string key = tune_key(ctx);
    TuneCfg best = current_tuning(ctx);
    float best_ms = 0.f, dflt_ms = 0.f;
    int n_timed = 0, st = QTP_OK;
    bool cached = false;
    if (!(flags & 1)) {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            best = it->second;
            cached = true;
        }
    }
    if (!cached) {
        const TuneCfg dflt = current_tuning(ctx);
        st = tune_time(ctx, 5, &dflt_ms);
        best_ms = dflt_ms;
        for (const TuneCfg &c : tune_candidates(ctx)) {
            if (st != QTP_OK) break;
            apply_tuning(ctx, c);
            float ms = 0.f;
            st = tune_time(ctx, 5, &ms);
            ++n_timed;
            // a candidate has to beat the configuration in hand by TUNE_MARGIN,
            // so that timing noise between near-equal ones does not decide
            if (st == QTP_OK && ms < best_ms * (1.0f - TUNE_MARGIN)) {
                best_ms = ms;
                best = c;
            }
        }
        // then the real-space share of the second stream and the structure-
        // factor chunks per block, one at a time on the best so far
        for (int sp = 1; sp <= 4 + 5 && st == QTP_OK; ++sp) {
            if (sp > 4 && !best.kmf) break;
            TuneCfg c = best;
            if (sp <= 4) c.rsplit = sp;
            else c.cpb = 1 << (sp - 5);
            apply_tuning(ctx, c);
            float ms = 0.f;
            st = tune_time(ctx, 5, &ms);
            ++n_timed;
            if (st == QTP_OK && ms < best_ms * (1.0f - TUNE_MARGIN)) {
                best_ms = ms;
                best = c;
            }
        }
        if (st == QTP_OK) {
            std::lock_guard<std::mutex> g(g_tune_mu);
            g_tune_cache[key] = best;
        } else {
            best = dflt;
        }
    }
    best.fused = ctx->prop_fused;   // not timed here: the caller's choice stays
    apply_tuning(ctx, best);
    ctx->timing = timing;
    CHK(st);
    CHK(tune_refresh(ctx));
    CHK(read_status(ctx));
    if (report) {
        report[0] = best_ms;
        report[1] = dflt_ms;
        report[2] = best.lpm;
        report[3] = best.kmf;
        report[4] = best.carg;
        report[5] = best.small;
        report[6] = best.usplit;
        report[7] = n_timed;
        report[8] = best.rsw;
        report[9] = best.rsplit;
        report[10] = best.cpb;
    }
    return QTP_OK;
}

// ---- Verlet-list build -------------------------------------------------------------

int qtp_set_lists(qtp_ctx *ctx, int method, int rebuild_every)
{
    if (!ctx) return QTP_ERR_ARG;
    if (method < QTP_LIST_AUTO || method > QTP_LIST_CELLS || rebuild_every < 0) {
        ctx->err = "lists: method QTP_LIST_AUTO / ALLPAIRS / CELLS and rebuild_every >= 0";
        return QTP_ERR_ARG;
    }
    if (method == QTP_LIST_CELLS && !ctx->cells_ok) {
        ctx->err = "lists: the box holds fewer than 3 cells of r_skin a side (or the system "
                   "exceeds 65536 molecules / 12288 cells)";
        return QTP_ERR_ARG;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->list_method = method;
    ctx->rebuild_every = rebuild_every;
    invalidate_graphs(ctx);
    return QTP_OK;
}

int qtp_get_list_info(const qtp_ctx *ctx, int info[5])
{
    if (!ctx || !info) return QTP_ERR_ARG;
    info[0] = use_cells(ctx) ? 1 : 0;
    for (int d = 0; d < 3; ++d) info[1 + d] = ctx->cells_ok ? ctx->cg.nc[d] : 0;
    info[4] = ctx->cells_ok ? ctx->cg.S : 0;
    return QTP_OK;
}

// ---- centroid dynamics, thermostats, mean forces, g(r) ------------------------------

int qtp_set_thermostat(qtp_ctx *ctx, int kind, double tau_fs)
{
    if (!ctx) return QTP_ERR_ARG;
    if ((kind != QTP_THERMO_PILE_G && kind != QTP_THERMO_PILE_L) || !(tau_fs >= 0.0) ||
        !std::isfinite(tau_fs) || (kind == QTP_THERMO_PILE_L && tau_fs == 0.0)) {
        ctx->err = "thermostat: PILE_G with tau >= 0 or PILE_L with tau > 0";
        return QTP_ERR_ARG;
    }
    if (ctx->P == 1 && kind != QTP_THERMO_PILE_L) {
        ctx->err = "classical resident steps take the Langevin (PILE_L) thermostat only";
        return QTP_ERR_ARG;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->cthermo = kind;
    ctx->tau_fs = tau_fs;
    invalidate_graphs(ctx);
    return QTP_OK;
}

int qtp_set_pimd_mode(qtp_ctx *ctx, int mode, double omega_ad)
{
    if (!ctx) return QTP_ERR_ARG;
    if (mode != QTP_MODE_PIMD && mode != QTP_MODE_PACMD && mode != QTP_MODE_CENTROID_FIXED) {
        ctx->err = "pimd mode: QTP_MODE_PIMD, QTP_MODE_PACMD or QTP_MODE_CENTROID_FIXED";
        return QTP_ERR_ARG;
    }
    if (ctx->P < 2 && mode != QTP_MODE_PIMD) {
        ctx->err = "centroid dynamics modes need n_beads >= 2";
        return QTP_ERR_ARG;
    }
    double W = 0.0;
    if (mode == QTP_MODE_PACMD) {
        // Hone, Rossky, Voth, J. Chem. Phys. 124, 154103 (2006): W = P^(P/(P-1)) / (beta hbar)
        const double P = ctx->P;
        W = omega_ad > 0.0 ? omega_ad : std::pow(P, P / (P - 1.0)) / ctx->drv.beta;
        if (!std::isfinite(W)) {
            ctx->err = "adiabatic frequency not finite";
            return QTP_ERR_ARG;
        }
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->pimd_mode = mode;
    ctx->omega_ad = W;
    CHK(upload_modes(ctx));
    invalidate_graphs(ctx);
    return QTP_OK;
}

int qtp_centroid_force(qtp_ctx *ctx, int on)
{
    if (!ctx) return QTP_ERR_ARG;
    if (ctx->P < 2) {
        ctx->err = "centroid mean force needs n_beads >= 2";
        return QTP_ERR_ARG;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (on) {
        HIPCHK(hipMemset(ctx->cmf_acc, 0, 3 * ctx->NAP * sizeof(double)));
        HIPCHK(hipMemset(ctx->cmf_n, 0, sizeof(unsigned long long)));
    }
    if ((on != 0) != ctx->cmf_on) invalidate_graphs(ctx);
    ctx->cmf_on = on != 0;
    return QTP_OK;
}

int qtp_get_centroid_force(qtp_ctx *ctx, double *mean_force, double *centroid, int *n_samples)
{
    if (!ctx || !mean_force) return QTP_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    unsigned long long n = 0;
    HIPCHK(hipMemcpy(&n, ctx->cmf_n, sizeof(n), hipMemcpyDeviceToHost));
    std::vector<double> h(3 * (size_t)ctx->NAP);
    HIPCHK(hipMemcpy(h.data(), ctx->cmf_acc, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int a = 0; a < ctx->n_atoms; ++a)
        for (int d = 0; d < 3; ++d)
            mean_force[3 * a + d] = n ? h[(size_t)d * ctx->NAP + a] / (double)n : 0.0;
    if (centroid) {
        k_centroid<<<(3 * ctx->NAP + 255) / 256, 256, 0, ctx->stream>>>(ctx->P, ctx->NAP, ctx->q,
                                                                       ctx->qc);
        HIPCHK(hipGetLastError());
        CHK(download_planar(ctx, ctx->qc, 1, centroid));
    }
    if (n_samples) *n_samples = (int)n;
    return QTP_OK;
}

int qtp_rdf_reset(qtp_ctx *ctx, double r_max, int n_bins)
{
    if (!ctx || !(r_max > 0.0) || n_bins < 1 || n_bins > (1 << 20)) return QTP_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (n_bins != ctx->rdf_bins) {
        if (ctx->rdf_oo) hipFree(ctx->rdf_oo);
        if (ctx->rdf_oh) hipFree(ctx->rdf_oh);
        ctx->rdf_oo = ctx->rdf_oh = nullptr;
        ctx->rdf_bins = 0;
        CHK(dalloc(ctx, &ctx->rdf_oo, n_bins));
        CHK(dalloc(ctx, &ctx->rdf_oh, n_bins));
        ctx->rdf_bins = n_bins;
    }
    HIPCHK(hipMemset(ctx->rdf_oo, 0, n_bins * sizeof(unsigned int)));
    HIPCHK(hipMemset(ctx->rdf_oh, 0, n_bins * sizeof(unsigned int)));
    ctx->rdf_dr = r_max / n_bins;
    ctx->rdf_n = 0;
    return QTP_OK;
}

int qtp_rdf_sample(qtp_ctx *ctx)
{
    if (!ctx) return QTP_ERR_ARG;
    if (!ctx->rdf_bins || !ctx->state_set) {
        ctx->err = "qtp_rdf_sample: no histogram (qtp_rdf_reset) or no resident state";
        return QTP_ERR_STATE;
    }
    const double *x = ctx->q;
    if (ctx->P > 1) {
        k_centroid<<<(3 * ctx->NAP + 255) / 256, 256, 0, ctx->stream>>>(ctx->P, ctx->NAP, ctx->q,
                                                                       ctx->qc);
        x = ctx->qc;
    }
    k_rdf<<<(ctx->n_mols + 3) / 4, 256, 0, ctx->stream>>>(ctx->n_mols, ctx->NAP, x, ctx->box,
                                                          1.0 / ctx->rdf_dr, ctx->rdf_bins,
                                                          ctx->rdf_oo, ctx->rdf_oh);
    HIPCHK(hipGetLastError());
    ctx->rdf_n += 1;
    return QTP_OK;
}

int qtp_rdf_get(qtp_ctx *ctx, double *oo, double *oh, int *n_samples)
{
    if (!ctx || !oo || !oh) return QTP_ERR_ARG;
    if (!ctx->rdf_bins) {
        ctx->err = "qtp_rdf_get: no histogram (qtp_rdf_reset)";
        return QTP_ERR_STATE;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    std::vector<unsigned int> a(ctx->rdf_bins), b(ctx->rdf_bins);
    HIPCHK(hipMemcpy(a.data(), ctx->rdf_oo, a.size() * sizeof(unsigned int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(b.data(), ctx->rdf_oh, b.size() * sizeof(unsigned int), hipMemcpyDeviceToHost));
    for (int t = 0; t < ctx->rdf_bins; ++t) {
        oo[t] = a[t];
        oh[t] = b[t];
    }
    if (n_samples) *n_samples = ctx->rdf_n;
    return QTP_OK;
}

int qtp_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    if (!ctr || !key || !out) return QTP_ERR_ARG;
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox4x32_10(c, key[0], key[1]);
    for (int i = 0; i < 4; ++i) out[i] = c[i];
    return QTP_OK;
}

}  // extern "C"
