// qtp_host_nn.hpp -- host side of the learned centroid-force correction
// (device code in qtp_nn.hpp): the model object, batched evaluation,
// force-matching training, and the coupling into a context's classical
// resident step.  Included once by qtip4pf.hip (one translation unit).
#pragma once

struct qtp_nn {
    std::string err;
    int n_mols = 0, n_atoms = 0, NAP = 0, max_frames = 0;
    Box box{};
    hipStream_t stream = nullptr;
    // model: FP32 master parameters, Adam moments, the device copies
    float *theta = nullptr, *am = nullptr, *av = nullptr;
    __bf16 *w1 = nullptr, *w1t = nullptr, *w2 = nullptr, *w2t = nullptr;
    float *b1 = nullptr, *b2 = nullptr, *w3 = nullptr, *b3 = nullptr;
    float *mean = nullptr, *stdv = nullptr;
    double e_scale = 1.0;
    int adam_t = 0;
    bool model_set = false;
    // per-batch work space (max_frames frames)
    double *pos = nullptr, *tgt = nullptr, *fwork = nullptr, *dsum = nullptr;
    float *gt = nullptr, *ut = nullptr, *dEdG = nullptr, *Eatom = nullptr, *res = nullptr;
    float *sq = nullptr, *graw = nullptr, *part = nullptr, *tp = nullptr, *grad = nullptr;
    int *nbr = nullptr, *nn = nullptr, *nba = nullptr, *nna = nullptr, *status = nullptr;
    float4 *dang = nullptr, *dself = nullptr;   // angular force pieces per centre
    float4 *P4 = nullptr;                         // positions (Angstrom, wrapped)
    float *Etile = nullptr;                       // energy per inference tile
    __bf16 *inter = nullptr;   // the eight training intermediates, both elements
    size_t inter_elems = 0;
    // training data (device resident)
    double *dpos = nullptr, *dtgt = nullptr;
    int n_data = 0;
};

namespace {

using namespace qtp::nn;

#define NNCHK(x)                                                                     \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            nn->err = std::string(#x) + ": " + hipGetErrorString(e_);                \
            return QTP_ERR_HIP;                                                      \
        }                                                                            \
    } while (0)

template <class T>
static int nalloc(qtp_nn *nn, T **p, size_t n)
{
    NNCHK(hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
    NNCHK(hipMemset(*p, 0, std::max<size_t>(n, 1) * sizeof(T)));
    return QTP_OK;
}

// Launch geometry of a batch of F frames: per-atom kernels (one block per
// atom), inference tiles of 16 rows per element, training tiles padded to
// TRAIN_WAVES per block, weight-gradient chunks.
struct NNBatch {
    int F, nblk_atom, to, th, nbo, nbh, tiles_o, tiles_h, rows_o, rows_h, nco, nch;
};

static NNBatch nn_batch(const qtp_nn *nn, int F)
{
    NNBatch b;
    b.F = F;
    b.nblk_atom = F * nn->n_atoms;   // one block of AT threads per atom
    b.to = (F * nn->n_mols + 15) / 16;
    b.th = (2 * F * nn->n_mols + 15) / 16;
    b.nbo = (b.to + TRAIN_WAVES - 1) / TRAIN_WAVES;
    b.nbh = (b.th + TRAIN_WAVES - 1) / TRAIN_WAVES;
    b.tiles_o = b.nbo * TRAIN_WAVES;
    b.tiles_h = b.nbh * TRAIN_WAVES;
    b.rows_o = ((b.tiles_o * 16 + WG_ROWS - 1) / WG_ROWS) * WG_ROWS;
    b.rows_h = ((b.tiles_h * 16 + WG_ROWS - 1) / WG_ROWS) * WG_ROWS;
    b.nco = b.rows_o / WG_ROWS;
    b.nch = b.rows_h / WG_ROWS;
    return b;
}

static NNGeom nn_geom(const qtp_nn *nn, int F)
{
    NNGeom g;
    g.n_atoms = nn->n_atoms;
    g.n_mols = nn->n_mols;
    g.n_frames = F;
    g.NAP = nn->NAP;
    g.box = nn->box;
    for (int k = 0; k < 3; ++k) {
        g.Lf[k] = (float)(nn->box.L[k] * ANGD);
        g.iLf[k] = 1.0f / g.Lf[k];
    }
    return g;
}

static NNW nn_weights(const qtp_nn *nn)
{
    NNW W;
    W.w1 = nn->w1;
    W.w1t = nn->w1t;
    W.w2 = nn->w2;
    W.w2t = nn->w2t;
    W.b1 = nn->b1;
    W.b2 = nn->b2;
    W.w3 = nn->w3;
    W.b3 = nn->b3;
    W.e_scale = (float)nn->e_scale;
    return W;
}

static NNDev nn_dev(const qtp_nn *nn)
{
    NNDev D;
    D.w1 = nn->w1;
    D.w1t = nn->w1t;
    D.w2 = nn->w2;
    D.w2t = nn->w2t;
    D.b1 = nn->b1;
    D.b2 = nn->b2;
    D.w3 = nn->w3;
    D.b3 = nn->b3;
    return D;
}

static NNTrainBuf nn_trainbuf(qtp_nn *nn, const NNBatch &b)
{
    NNTrainBuf T;
    const int rows[2] = {b.rows_o, b.rows_h};
    __bf16 *p = nn->inter;
    __bf16 **slot[8][2] = {{&T.d2t[0], &T.d2t[1]}, {&T.d2p[0], &T.d2p[1]},
                           {&T.a1t[0], &T.a1t[1]}, {&T.a1p[0], &T.a1p[1]},
                           {&T.d1t[0], &T.d1t[1]}, {&T.d1p[0], &T.d1p[1]},
                           {&T.a0t[0], &T.a0t[1]}, {&T.a0p[0], &T.a0p[1]}};
    for (int e = 0; e < 2; ++e)
        for (int s = 0; s < 8; ++s) {
            *slot[s][e] = p;
            p += (size_t)(s < 6 ? NH : KIN) * rows[e];
        }
    T.rows[0] = rows[0];
    T.rows[1] = rows[1];
    T.tp = nn->tp;
    return T;
}

static size_t nn_inter_elems(const NNBatch &b)
{
    return (size_t)(6 * NH + 2 * KIN) * (b.rows_o + b.rows_h);
}

// Positions, lists, descriptors, energies, dE/dG and the angular force
// pieces of F frames at pos (planar FP64, device), the atoms' status bits
// into status; Etile (may be null) receives the tile energy sums.
static int nn_forward(qtp_nn *nn, hipStream_t s, const double *pos, int F, const NNBatch &b,
                      int *status, float *graw = nullptr, float *Etile = nullptr)
{
    const NNGeom g = nn_geom(nn, F);
    k_nn_pos<<<(F * nn->n_atoms + 255) / 256, 256, 0, s>>>(g, pos, nn->P4);
    k_nn_desc<<<b.nblk_atom, AT, 0, s>>>(g, nn->P4, nn->mean, nn->stdv, nn->gt, nn->nbr, nn->nn,
                                         nn->nba, nn->nna, graw, status);
    k_nn_mlp<<<b.to + b.th, 64 * MLP_WAVES, 0, s>>>(g, nn_weights(nn), b.to, nn->gt, nn->stdv,
                                                    nn->dEdG, nn->Eatom, Etile);
    k_nn_dang<<<b.nblk_atom, AT, 0, s>>>(g, nn->P4, nn->nba, nn->nna, nn->dEdG, nn->dang,
                                         nn->dself);
    NNCHK(hipGetLastError());
    return QTP_OK;
}

static int nn_read_status(qtp_nn *nn)
{
    int st = 0;
    NNCHK(hipMemcpy(&st, nn->status, sizeof(int), hipMemcpyDeviceToHost));
    if (st & 4) {
        nn->err = "an atom has more neighbours than the descriptor lists hold (192 within 6 A, "
                  "64 within 4 A)";
        return QTP_ERR_STATE;
    }
    return QTP_OK;
}

static int nn_upload_frames(qtp_nn *nn, double *dst, const double *ref, int F)
{
    const long N3 = 3L * nn->NAP;
    std::vector<double> h((size_t)F * N3, 0.0);
    for (int f = 0; f < F; ++f)
        for (int a = 0; a < nn->n_atoms; ++a)
            for (int d = 0; d < 3; ++d)
                h[f * N3 + (long)d * nn->NAP + a] = ref[((long)f * nn->n_atoms + a) * 3 + d];
    NNCHK(hipMemcpy(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return QTP_OK;
}

// One step of force matching on frames [f0, f0 + F) of the resident data:
// mode 1 = Adam update, 0 = gradient into nn->grad.  *loss_dev gets the
// batch loss (mean squared force residual) before the step.
static int nn_train_batch(qtp_nn *nn, hipStream_t s, int f0, int F, int mode, double lr,
                          double *loss_dev)
{
    const NNBatch b = nn_batch(nn, F);
    const NNGeom g = nn_geom(nn, F);
    const long N3 = 3L * nn->NAP;
    const double *pos = nn->dpos + (long)f0 * N3, *tgt = nn->dtgt + (long)f0 * N3;
    CHK(nn_forward(nn, s, pos, F, b, nn->status));
    k_nn_force<1><<<b.nblk_atom, AT, 0, s>>>(g, nn->P4, nn->nbr, nn->nn, nn->nba, nn->nna,
                                             nn->dEdG, nn->dang, nn->dself, nullptr, tgt, nn->res,
                                             nn->sq);
    const long ncomp = 3L * F * nn->n_atoms;
    k_nn_sum<<<1, 1024, 0, s>>>(F * nn->n_atoms, nn->sq, 1.0 / (double)ncomp, 0, loss_dev);
    k_nn_tan<<<b.nblk_atom, AT, 0, s>>>(g, nn->P4, nn->nbr, nn->nn, nn->nba, nn->nna, nn->res,
                                        nn->stdv, nn->ut);
    NNTrainBuf T = nn_trainbuf(nn, b);
    NNCHK(hipMemsetAsync(nn->inter, 0, nn_inter_elems(b) * sizeof(__bf16), s));
    k_nn_train_fwd<<<b.nbo + b.nbh, 64 * MLP_WAVES, TRAIN_LDS, s>>>(g, nn_weights(nn), b.nbo, nn->gt,
                                                                   nn->ut, T);
    k_nn_wgrad<<<b.nco + b.nch, 64 * WG_WAVES, 0, s>>>(T, b.nco, nn->part);
    AdamArgs ad{};
    if (mode == 1) {
        nn->adam_t += 1;
        ad.lr = (float)lr;
        ad.beta1 = 0.9f;
        ad.beta2 = 0.999f;
        ad.eps = 1e-8f;
        ad.c1 = (float)(1.0 - std::pow(0.9, nn->adam_t));
        ad.c2 = (float)(1.0 - std::pow(0.999, nn->adam_t));
    }
    k_nn_update<<<(2 * NPE + 255) / 256, 256, 0, s>>>(mode, nn->part, b.nco, b.nch, nn->tp,
                                                      b.tiles_o, b.tiles_h, -2.0 / (double)ncomp,
                                                      nn->theta, nn->am, nn->av, ad, nn->grad,
                                                      nn_dev(nn));
    NNCHK(hipGetLastError());
    return QTP_OK;
}

}  // namespace

// The correction's forces on the resident classical state into ctx->fnn and
// its energy as tile sums in nn->Etile (both read by k_vv_b), on stream s.
static int run_nn_forces(qtp_ctx *ctx, hipStream_t s)
{
    qtp_nn *nn = ctx->nn;
    const NNBatch b = nn_batch(nn, 1);
    const NNGeom g = nn_geom(nn, 1);
    {
        TimeScope ts(ctx, "nn_desc");
        k_nn_pos<<<(nn->n_atoms + 255) / 256, 256, 0, s>>>(g, ctx->q, nn->P4);
        k_nn_desc<<<b.nblk_atom, AT, 0, s>>>(g, nn->P4, nn->mean, nn->stdv, nn->gt, nn->nbr, nn->nn,
                                             nn->nba, nn->nna, nullptr, ctx->status);
    }
    {
        TimeScope ts(ctx, "nn_mlp");
        k_nn_mlp<<<b.to + b.th, 64 * MLP_WAVES, 0, s>>>(g, nn_weights(nn), b.to, nn->gt, nn->stdv,
                                                        nn->dEdG, nullptr, nn->Etile);
    }
    {
        TimeScope ts(ctx, "nn_dang");
        k_nn_dang<<<b.nblk_atom, AT, 0, s>>>(g, nn->P4, nn->nba, nn->nna, nn->dEdG, nn->dang,
                                             nn->dself);
    }
    {
        TimeScope ts(ctx, "nn_force");
        k_nn_force<0><<<b.nblk_atom, AT, 0, s>>>(g, nn->P4, nn->nbr, nn->nn, nn->nba, nn->nna,
                                                 nn->dEdG, nn->dang, nn->dself, ctx->fnn, nullptr,
                                                 nullptr, nullptr);
    }
    HIPCHK(hipGetLastError());
    return QTP_OK;
}

static int nn_tiles(const qtp_nn *nn)
{
    const NNBatch b = nn_batch(nn, 1);
    return b.to + b.th;
}
