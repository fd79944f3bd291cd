// qtp_host_ctx.hpp -- host side of the gfx950 qTIP4P/F step: the context
// (one system: buffers, streams, graphs, derived parameters), error and
// allocation helpers, per-kernel timing.
// Included once, in order, by qtip4pf.hip (one translation unit).
#pragma once

namespace {


struct Timer {
    double ms = 0.0;
    int n = 0;
};

}  // namespace

struct qtp_nn;

struct qtp_ctx {
    qtp_params prm{};
    qtp_derived drv{};
    std::string err;
    hipStream_t stream = nullptr;
    // second stream for the k-space half of a force phase (fork/join by events)
    hipStream_t s2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // third stream for the learned correction of classical steps
    hipStream_t s3 = nullptr;
    hipEvent_t ev_nn_fork = nullptr, ev_nn_join = nullptr;

    int P = 0, n_mols = 0, n_atoms = 0, n_dof = 0, NAP = 0, NMP = 0, nblk = 0, list_size = 0;
    long fcap = 0;
    Box box{};
    FF ff{};
    Ewald ew{};
    double eself = 0.0;  // sum_i 0.5 * (2 kappa/sqrt(pi)) q_i^2 (forces.cpp:622-624)
    int G = 0;           // k groups
    int ntile = 0;       // real-space tiles per evaluation
    int nmb = 0;         // LJ/intra blocks per evaluation
    int lj_lpm = 16;     // lanes per molecule in k_ljintra
    bool rs_cull = false;  // k_realspace skips tiles beyond r_c (boxes >= RS_CULL_MIN_BOX r_c)
    double4 *rs_sites = nullptr;  // CULL: [E][nblk * 64] sites (x, y, z, q)
    double *rs_box = nullptr;     // CULL: [E][nblk][RS_BOX] block and 8-site cluster bounding boxes
    unsigned char *rs_far = nullptr;  // CULL: [E][ntile] 1 = tile skipped this evaluation
    int AC = 64, nac = 0, ngc = 0, ngr = 0;
    int ep_stride = 0, ep_lj = 0, ep_real = 0, ep_k = 0, ep_spl = 0;
    // splined pair potentials (fast-CMD path): tables loaded, used by the
    // resident steps, used by the current call
    bool spl_loaded = false, spl_resident = false, spl_on = false;
    SplTab spl{};
    double *spl_coef = nullptr, *fspl = nullptr;
    int ptile = 0;       // propagation tiles (64 planar dofs each)
    size_t kforce_lds = 0, kforce_lds_split = 0, ksum_lds = 0, prop_lds = 0;
    int kf_chunk = KF_CHUNK, kf_chunk_split = KF_CHUNK;  // k_kforce record staging
    // split k_kforce in the SC phases when its P * ceil(n_atoms/63) blocks
    // would not fill two per CU (small systems); large ones have blocks enough
    bool kf_split = false, kf_split_ok = false;
    // k-space sums and forces on the matrix cores (k_ksum_mf, k_kforce_mf)
    // for k_max <= 23; the vector kernels above that, or when
    // QTP_KSPACE=vector is set at create
    bool kf_mf = false, kf_mf_ok = false;
    // launch choices the autotuner may set (qtp_set_tuning): k_realspace
    // coefficients in SGPRs (-1: for grids of RS_CARG_BLOCKS blocks or more),
    // k_kforce_mf 16-wave blocks (-1: below MF_SMALL_BLOCKS blocks), k_ksum_mf
    // blocks per chunk (0: several when a launch has under 128 chunks)
    int rs_carg = -1, mf_small = -1, ks_usplit = 0;
    int rs_waves = RS_WAVES;  // waves per k_realspace tile block (2 or 4)
    int rs_split = 0;         // eighths of the real-space tiles on the second stream (0-4)
    int ks_cpb = 0;           // structure-factor chunks per k_ksum_mf block (0: by launch size)
    // a list check the SC step leaves to the next force phase's second stream
    bool lists_defer = false, lists_heat = false;
    const unsigned long long *lists_ctr = nullptr;
    // propagation as one LDS-staged kernel per step (k_prop_tile) where the
    // thermostat allows it: -1 for systems of PT_MIN_TILES tiles or more,
    // 1 always, 0 never (k_prop_fwd + k_prop_inv)
    int prop_fused = -1;
    int ep_cap = 0;   // allocated stride of the energy partials
    int mf_Mp = 0, mf_nks = 0, mf_nrt = 0, mf_nrt_small = 0, ks_nunit = 0, ks_ncol4 = 0;
    size_t kforce_mf_lds = 0, kforce_mf_lds_small = 0, ksum_mf_lds = 0;

    // device state (planar SoA)
    double *q = nullptr, *p = nullptr, *F = nullptr, *fph = nullptr, *disp = nullptr;
    double *m_atom = nullptr, *sqm_atom = nullptr, *charge = nullptr;
    int *point = nullptr, *list = nullptr, *fpoint = nullptr, *flist = nullptr, *cnt = nullptr;
    int *flag = nullptr, *status = nullptr;
    // cell-binned list build (k_cell_build, k_nl_cell): the grid when the box
    // holds one (cells_ok), the method (QTP_LIST_AUTO / ALLPAIRS / CELLS) and
    // the forced rebuild period of resident steps (0: displacement test only)
    CellGrid cg{};
    bool cells_ok = false;
    int list_method = 0, rebuild_every = 0;
    int *cell_of = nullptr, *crank = nullptr, *cstart = nullptr, *member = nullptr;
    double4 *cell_pos = nullptr;  // O positions in member order
    // the count pass's row bitmaps for the fill pass (when they take at most
    // NL_BM_MAX_BYTES; otherwise the fill pass tests the candidates again)
    unsigned long long *nl_bm = nullptr;
    double *nreb = nullptr;
    // force-phase scratch (E <= max(P,1))
    double *qpm = nullptr, *vpart = nullptr, *vout = nullptr, *rpart = nullptr, *spart = nullptr, *sprime = nullptr;
    double *flj = nullptr, *fpm = nullptr, *Ve = nullptr, *epart = nullptr;
    double *fk2 = nullptr;  // second half of split k-space forces (SC phases)
    double *wm = nullptr;   // k_kforce_mf coefficient matrices [E][4 nks][Mp]
    int *gcol = nullptr;    // column of each k-group in the matrix rows
    int4 *mf_rt = nullptr;  // k_kforce_mf row-tile table (MF_WAVES waves)
    int4 *mf_rt_small = nullptr;  // and for MF_WAVES_SMALL
    int4 *ks_col = nullptr; // k_ksum_mf column table
    int2 *ks_unit = nullptr;
    double *scal = nullptr, *out = nullptr, *heat = nullptr, *heat_step = nullptr;
    unsigned long long *step_ctr = nullptr;
    // constants
    int4 *groups = nullptr;
    double *gcf = nullptr;
    double4 *grk = nullptr;
    int2 *tiles = nullptr;
    double *Cm = nullptr;
    double4 *mode = nullptr;
    double2 *ms = nullptr;   // per-mode propagation-mass factor {f, sqrt f}
    // centroid dynamics and thermostat of the resident steps: pimd_mode 0
    // PIMD, 1 partially adiabatic CMD at frequency omega_ad, 2 centroid held
    // fixed; cthermo 0 PILE-G (Bussi, time constant tau_fs; 0 = the
    // reference's pile_g), 1 PILE-L (Langevin, friction 1/tau_fs)
    int pimd_mode = 0, cthermo = 0;
    double omega_ad = 0.0, tau_fs = 0.0;
    // centroid mean force accumulated over resident steps (k_cmf_acc)
    bool cmf_on = false;
    double *cmf_acc = nullptr;
    unsigned long long *cmf_n = nullptr;
    // radial distribution functions of the (centroid) positions
    double *qc = nullptr;
    unsigned int *rdf_oo = nullptr, *rdf_oh = nullptr;
    int rdf_bins = 0, rdf_n = 0;
    double rdf_dr = 0.0;
    // propagation scratch
    double *nq = nullptr, *np = nullptr, *ppart = nullptr, *hpart = nullptr, *dev = nullptr;
    double *obs = nullptr;
    long dev_cap = 0;
    uint64_t seed = 0;

    // host copies
    std::vector<double> h_mass_atom, h_charge;
    bool state_set = false;
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // [is_nvt]
    // resident observables: result, sample trace, and the graphs that record it
    double *obs_out = nullptr, *trace = nullptr;
    int *ntrace = nullptr;
    int trace_cap = 0, obs_every = 0;
    hipGraphExec_t graph_obs[2] = {nullptr, nullptr};
    // learned centroid-force correction added to classical resident steps
    qtp_nn *nn = nullptr;
    double *fnn = nullptr;   // its forces (planar)
    bool timing = false;
    std::map<std::string, Timer> timers;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> evpool;
    size_t evnext = 0;
};

// ---------------------------------------------------------------------------
#define HIPCHK(x)                                                                    \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            ctx->err = std::string(#x) + ": " + hipGetErrorString(e_);               \
            return QTP_ERR_HIP;                                                      \
        }                                                                            \
    } while (0)

#define CHK(x)                                                                       \
    do {                                                                             \
        int s_ = (x);                                                                \
        if (s_ != QTP_OK) return s_;                                                 \
    } while (0)

template <class T>
static int dalloc(qtp_ctx *ctx, T **p, size_t n)
{
    HIPCHK(hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
    HIPCHK(hipMemset(*p, 0, std::max<size_t>(n, 1) * sizeof(T)));
    return QTP_OK;
}

// Timed launch bracket (only when timing is enabled; never inside a graph).
static hipEvent_t ev_get(qtp_ctx *ctx)
{
    if (ctx->evnext >= ctx->evpool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->evpool.push_back(e);
    }
    return ctx->evpool[ctx->evnext++];
}

struct TimeScope {
    qtp_ctx *ctx;
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
    TimeScope(qtp_ctx *c, const char *n) : ctx(c), name(n)
    {
        if (ctx->timing) {
            a = ev_get(ctx);
            b = ev_get(ctx);
            if (a) hipEventRecord(a, ctx->stream);
        }
    }
    ~TimeScope()
    {
        if (ctx->timing && a && b) {
            hipEventRecord(b, ctx->stream);
            ctx->pending.push_back({name, {a, b}});
        }
    }
};

static int flush_timing(qtp_ctx *ctx)
{
    if (ctx->pending.empty()) return QTP_OK;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (auto &pe : ctx->pending) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
        Timer &t = ctx->timers[pe.first];
        t.ms += ms;
        t.n += 1;
    }
    ctx->pending.clear();
    ctx->evnext = 0;
    return QTP_OK;
}

