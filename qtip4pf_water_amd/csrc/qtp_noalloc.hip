// qtp_noalloc.hip -- the centroid-virial kinetic energy estimator for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops): qtp_get_kcv_noalloc in include/qtip4pf.h.
//
// Its block partials live in `scratch`, a fixed device array of this
// translation unit.  Host code never passes the array's name to a kernel (on
// the host that name is only a stand-in for the symbol): the device address
// is looked up once per device with hipGetSymbolAddress and kept in
// g_scratch[device].  Every call on one GPU shares that GPU's copy, so calls
// that overlap on one GPU race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/qtip4pf.h"
#include "qtp_common.hpp"

namespace {

constexpr int KCV_THREADS = 256;
// at most this many 256-dof blocks: n_dof <= 256 * 8192 = 2,097,152 (233,016 molecules)
constexpr int SCRATCH_DOUBLES = 8192;
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

std::mutex g_mu;
double *g_scratch[MAX_DEVICES] = {};
thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

// Device address of this device's `scratch` (looked up on first use).
int scratch_ptr(double **out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES)
        return fail(QTP_ERR_HIP, "qtp_get_kcv_noalloc: no current HIP device");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(QTP_ERR_HIP, std::string("qtp_get_kcv_noalloc: hipGetSymbolAddress: ") +
                                         hipGetErrorString(e));
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return QTP_OK;
}

// path_integrals.cpp:29-48 on the reference layout x[b*n_dof + j]: per dof,
// the centroid sum over beads divided by float(P) (path_integrals.cpp:21),
// then -(q_bj - qc_j) f_bj over even beads.  One partial per 256-dof block.
__global__ __launch_bounds__(KCV_THREADS) void k_kcv_ref(int P, int n_dof,
                                                         const double *__restrict__ q,
                                                         const double *__restrict__ f,
                                                         double *__restrict__ part)
{
    __shared__ double sh[16];
    const int j = blockIdx.x * KCV_THREADS + threadIdx.x;
    double acc = 0.0;
    if (j < n_dof) {
        double qc = 0.0;
        for (int b = 0; b < P; ++b) qc += q[(long)b * n_dof + j];
        qc /= (double)(float)P;
        for (int b = 0; b < P; b += 2) acc -= (q[(long)b * n_dof + j] - qc) * f[(long)b * n_dof + j];
    }
    acc = qtp::block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// K_cv = n_dof / (2 beta) + (1/P) sum of the nb partials (fixed order).
__global__ __launch_bounds__(KCV_THREADS) void k_kcv_fin(int P, int n_dof, int nb, double beta,
                                                         const double *__restrict__ part,
                                                         double *__restrict__ kcv)
{
    __shared__ double sh[16];
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += KCV_THREADS) s += part[i];
    s = qtp::block_sum(s, sh);
    if (threadIdx.x == 0) kcv[0] = 0.5 * n_dof / beta + s / P;
}

}  // namespace

const char *qtp_noalloc_last_error() { return g_err.c_str(); }

extern "C" int qtp_get_kcv_noalloc(int n_beads, int n_atoms, double beta, const double *q,
                                   const double *fphys, double *kcv, void *stream)
{
    if (n_beads < 1 || n_atoms < 1 || !(beta > 0.0) || !q || !fphys || !kcv)
        return fail(QTP_ERR_ARG, "qtp_get_kcv_noalloc: bad argument");
    const long n_dof = 3L * n_atoms;
    const long nb = (n_dof + KCV_THREADS - 1) / KCV_THREADS;
    if (nb > SCRATCH_DOUBLES)
        return fail(QTP_ERR_ARG, "qtp_get_kcv_noalloc: n_atoms = " + std::to_string(n_atoms) +
                                     " exceeds the fixed scratch (at most " +
                                     std::to_string(SCRATCH_DOUBLES * KCV_THREADS / 3) + " atoms)");
    double *part = nullptr;
    const int st = scratch_ptr(&part);
    if (st != QTP_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // every partial is overwritten by this call: nothing to clear
    k_kcv_ref<<<(int)nb, KCV_THREADS, 0, s>>>(n_beads, (int)n_dof, q, fphys, part);
    k_kcv_fin<<<1, KCV_THREADS, 0, s>>>(n_beads, (int)n_dof, (int)nb, beta, part, kcv);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(QTP_ERR_HIP, std::string("qtp_get_kcv_noalloc: launch: ") + hipGetErrorString(e));
    return QTP_OK;
}
