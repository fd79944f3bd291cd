// qtp_host_tune.hpp -- launch-configuration autotuner (host side; no
// counterpart in the reference, which has one CPU code path).
//
// The tunable launch choices of a context (lanes per molecule of
// k_ljintra, matrix-core or vector k-space kernels, k_realspace coefficients
// as literals or in scalar registers, two or four waves per tile and the
// share of its tiles run on the second stream, 8- or 16-wave k_kforce_mf
// blocks, k_ksum_mf blocks per site chunk and site chunks per block) are timed on the context's own resident
// state: one force evaluation of every bead (both Suzuki-Chin phases, or the
// classical phase) captured as a graph, as the resident steps run it, and
// replayed between HIP events on the context's stream; median of several
// runs of replays.  Every
// candidate computes the same forces up to summation order.  The fastest is
// kept and cached per system for the process.
#pragma once

#include <mutex>

namespace {

constexpr float TUNE_MARGIN = 0.02f;

struct TuneCfg {
    int lpm = 16, kmf = 0, carg = -1, small = -1, usplit = 0, fused = -1, rsw = RS_WAVES, rsplit = 0;
    int cpb = 0;
};

TuneCfg current_tuning(const qtp_ctx *ctx)
{
    TuneCfg c;
    c.lpm = ctx->lj_lpm;
    c.kmf = ctx->kf_mf ? 1 : 0;
    c.carg = ctx->rs_carg;
    c.small = ctx->mf_small;
    c.usplit = ctx->ks_usplit;
    c.fused = ctx->prop_fused;
    c.rsw = ctx->rs_waves;
    c.rsplit = ctx->rs_split;
    c.cpb = ctx->ks_cpb;
    return c;
}

bool tuning_valid(const qtp_ctx *ctx, const TuneCfg &c)
{
    if (c.lpm != 16 && c.lpm != 64) return false;
    if (c.kmf != 0 && c.kmf != 1) return false;
    if (c.kmf && !ctx->kf_mf_ok) return false;
    if (c.carg < -1 || c.carg > 1 || c.small < -1 || c.small > 1) return false;
    if (c.small == 1 && ctx->kforce_mf_lds_small > 64 * 1024) return false;
    if (c.fused < -1 || c.fused > 1) return false;
    if (c.rsw != 2 && c.rsw != 4) return false;
    if (c.rsplit < 0 || c.rsplit > 4) return false;
    if (c.cpb != 0 && c.cpb != 1 && c.cpb != 2 && c.cpb != 4 && c.cpb != 8 && c.cpb != 16) return false;
    return c.usplit >= 0 && c.usplit <= 4;
}

void apply_tuning(qtp_ctx *ctx, const TuneCfg &c)
{
    set_lj_lanes(ctx, c.lpm);
    ctx->kf_mf = c.kmf != 0;
    ctx->kf_split = !ctx->kf_mf && ctx->kf_split_ok;
    ctx->rs_carg = c.carg;
    ctx->mf_small = c.small;
    ctx->ks_usplit = c.usplit;
    ctx->prop_fused = c.fused;
    ctx->rs_waves = c.rsw;
    ctx->rs_split = c.rsplit;
    ctx->ks_cpb = c.cpb;
    invalidate_graphs(ctx);
}

std::mutex g_tune_mu;
std::map<std::string, TuneCfg> g_tune_cache;

std::string tune_key(const qtp_ctx *ctx)
{
    char b[320];
    std::snprintf(b, sizeof b, "%d/%d/%.17g/%.17g/%.17g/%.17g/%.17g/%.17g/%d/%.17g/%d",
                  ctx->n_mols, ctx->P, ctx->box.L[0], ctx->box.L[1], ctx->box.L[2],
                  ctx->drv.kappa, ctx->drv.rcut, ctx->drv.k_cut, ctx->drv.k_max,
                  ctx->drv.lj_rcut, ctx->spl_resident ? 1 : 0);
    return b;
}

// The force phases of one resident step (no kick, no step count).
int tune_forces(qtp_ctx *ctx)
{
    if (ctx->P >= 2) return run_sc_forces(ctx, 0, nullptr);
    return run_phase(ctx, ctx->q, 1, 0, 0, ctx->fph, nullptr);
}

// Time (ms) of one force evaluation as the resident steps run it: the
// evaluation captured as a graph and replayed in runs between two events
// (up to TUNE_RUN replays a run, fewer for slow evaluations: a run lasts
// about TUNE_RUN_MS); the median over n runs after one untimed run.
constexpr int TUNE_RUN = 8;
constexpr float TUNE_RUN_MS = 2.0f;
int tune_time(qtp_ctx *ctx, int n, float *ms)
{
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int st = tune_forces(ctx);
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
    if (st != QTP_OK) {
        if (g) hipGraphDestroy(g);
        return st;
    }
    if (ec != hipSuccess) {
        ctx->err = std::string("autotune capture: ") + hipGetErrorString(ec);
        return QTP_ERR_HIP;
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess) {
        ctx->err = std::string("autotune instantiate: ") + hipGetErrorString(ei);
        return QTP_ERR_HIP;
    }
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) st = QTP_ERR_HIP;
    std::vector<float> t;
    int run = 1;  // the untimed first run is one replay; it sizes the others
    for (int i = 0; i < n + 1 && st == QTP_OK; ++i) {
        if (hipEventRecord(a, ctx->stream) != hipSuccess) st = QTP_ERR_HIP;
        for (int r = 0; r < run && st == QTP_OK; ++r)
            if (hipGraphLaunch(x, ctx->stream) != hipSuccess) st = QTP_ERR_HIP;
        if (st == QTP_OK && hipEventRecord(b, ctx->stream) != hipSuccess) st = QTP_ERR_HIP;
        if (st == QTP_OK && hipEventSynchronize(b) != hipSuccess) st = QTP_ERR_HIP;
        float e = 0.f;
        if (st == QTP_OK && hipEventElapsedTime(&e, a, b) != hipSuccess) st = QTP_ERR_HIP;
        if (i == 0)
            run = std::max(1, std::min(TUNE_RUN, (int)(TUNE_RUN_MS / std::max(e, 1e-3f))));
        else
            t.push_back(e / run);
    }
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    hipGraphExecDestroy(x);
    if (st != QTP_OK) {
        ctx->err = "autotune: HIP error while timing";
        return st;
    }
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return QTP_OK;
}

std::vector<TuneCfg> tune_candidates(const qtp_ctx *ctx)
{
    std::vector<TuneCfg> out;
    const int umax = std::min(4, std::max(1, ctx->ks_nunit / KS_WAVES));
    for (int rsw : {4, 2})
        for (int lpm : {16, 64})
            for (int kmf = ctx->kf_mf_ok ? 1 : 0; kmf >= 0; --kmf)
                for (int carg : ctx->rs_cull ? std::vector<int>{-1} : std::vector<int>{0, 1}) {
                    TuneCfg c;
                    c.lpm = lpm, c.kmf = kmf, c.carg = carg, c.fused = ctx->prop_fused, c.rsw = rsw;
                    if (!kmf) {
                        out.push_back(c);
                        continue;
                    }
                    const int smax = ctx->kforce_mf_lds_small <= 64 * 1024 ? 1 : 0;
                    for (int small = 0; small <= smax; ++small)
                        for (int u = 1; u <= umax; u *= 2) {
                            c.small = small, c.usplit = u;
                            out.push_back(c);
                        }
                }
    return out;
}

// The state's forces, energies and status under the configuration in force.
int tune_refresh(qtp_ctx *ctx)
{
    if (!ctx->state_set) return QTP_OK;
    if (ctx->P >= 2) CHK(run_sc_forces(ctx, 0, nullptr));
    else CHK(class_initial_forces(ctx));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return QTP_OK;
}

}  // namespace
