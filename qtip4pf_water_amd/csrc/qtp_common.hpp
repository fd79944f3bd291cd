// qtp_common.hpp -- gfx950 device code of the qTIP4P/F path-integral step:
// device constants and structs, reductions, site coordinates, the
// real-space special functions.
//
// Layout in HBM (per bead / per force evaluation, "planar SoA"):
//   x[(b*3 + d)*NAP + a]   a = 3*mol + site (O,H1,H2), NAP = n_atoms padded
//                          to a multiple of 64 (one wavefront of sites)
// Every kernel is FP64.  The translation unit is compiled with
// -ffp-contract=off: products and sums that the reference evaluates
// separately stay separate (bit-exact neighbour lists depend on it), and the
// hot loops use explicit fma() where rounding-level differences are allowed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qtp_erfcx.h"

namespace qtp {

constexpr int WAVE = 64;

struct Box {
    double L[3];
    double inv[3];
};

// qtip4pfparams.h:106-133 values needed on the device
struct FF {
    double gamma_vs, epslj4, epslj24, sigmalj6, sigmalj12, e_cut, r_cut_sq, r_skin_sq;
    double roheq, ar2, ar3, ar4fac, Dr, theta_eq, ktheta;
    double ar2x2, ar3x3, ar4x4;  // 2*ar2, 3*ar3, 4*ar4fac (forces.cpp:83 products)
    double half_ktheta;          // 0.5*ktheta
    double half1mg;              // 0.5*(1-gamma_vs)
};

struct Ewald {
    double kappa, kappa2, rcut_sq, pref2k;  // pref2k = 2 kappa / sqrt(pi)
    double bk[3];                           // 2 pi / L_d (reference pi)
    int K;                                  // k_max + 1
};

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide sum (blockDim multiple of 64, <= 1024); result valid in thread 0.
// Workgroups are dispatched to the 8 XCDs (each with its own L2) round
// robin: workgroup b runs on XCD b % 8.  This maps b to a logical index so
// that each XCD gets one contiguous eighth of the logical range: blocks that
// share data (same evaluation, same dof tile) then share an L2.  Identity
// when the grid is not a multiple of 8.
__device__ __forceinline__ int xcd_remap(int b, int n)
{
    if (n & 7) return b;
    return (b & 7) * (n >> 3) + (b >> 3);
}

__device__ __forceinline__ double block_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; ++i) r += sh[i];
    }
    return r;
}

// Minimum image as neighbour.cpp:46 / forces.cpp:503: IEEE divide, round half
// away from zero, separate multiply and subtract.
__device__ __forceinline__ double mimg_exact(double r, double L)
{
    return r - L * round(r / L);
}

// Minimum image for force kernels: r*(1/L) and round-to-even differ from the
// exact form only when r/L lies within an ulp of a half-integer, i.e. for
// separations of half a box, which are beyond every force cutoff (<= L/2).
__device__ __forceinline__ double mimg_fast(double r, double L, double invL)
{
    return fma(-L, rint(r * invL), r);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) and Box-Muller normals
// ---------------------------------------------------------------------------
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// One standard normal for (seed, step, stream, index).
__device__ inline double philox_normal(uint64_t seed, uint64_t step, uint32_t stream, uint32_t idx)
{
    uint32_t c[4] = {idx, (uint32_t)step, stream, (uint32_t)(step >> 32)};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double inv53 = 1.0 / 9007199254740992.0;  // 2^-53
    const double u1 = ((double)(((uint64_t)(c[0] >> 5) << 26) | (c[1] >> 6)) + 0.5) * inv53;
    const double u2 = ((double)(((uint64_t)(c[2] >> 5) << 26) | (c[3] >> 6)) + 0.5) * inv53;
    return sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

// ---------------------------------------------------------------------------
// Force evaluation.  A "phase" evaluates E independent single-bead forces
// (forces.cpp:468-630); eval e reads positions pos + e*pstride and bead
// lst_bead(e)'s Verlet list.  Energies are left as per-block partials
// epart[e][0 .. ep_stride) (LJ+intra blocks, real-space tiles, k-space
// blocks, spline blocks); eval_energy() sums the first ep_n of them (the
// spline blocks only when the splined potential is on) in that fixed order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lst_bead(int mode, int base, int e)
{
    return mode == 0 ? base + e : 2 * (e >> 1) + 1;
}

// Charge-site coordinate d of site a (a < n_atoms): the M-site for O slots,
// r_M = gamma r_O + (1-gamma)/2 (r_H1 + r_H2) (forces.cpp:558-564).
__device__ __forceinline__ double site_coord(const double *__restrict__ P, int NAP, int a, int d,
                                             double gamma, double half1mg)
{
    const double v = P[d * NAP + a];
    if (a % 3 != 0) return v;
    return gamma * v + half1mg * (P[d * NAP + a + 1] + P[d * NAP + a + 2]);
}

// Sum of n partials by one wavefront (lane-strided, then a fixed xor tree):
// deterministic, the same value in every lane.
__device__ __forceinline__ double wave_sum_array(const double *__restrict__ p, int n)
{
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    // unrolled so that sixteen loads are in flight; each lane still adds its
    // partials in index order (the same result as the plain loop).  4096 x 32:
    // k_sc1 172 -> 70 us, its wavefront summing ~21k partials per bead
#pragma unroll 16
    for (int k = lane; k < n; k += 64) s += p[k];
    return wave_sum(s);
}

// Energy of evaluation e: its partials (LJ+intra blocks, real-space tiles,
// k-space blocks) minus the self term.  Called by a whole wavefront.
__device__ __forceinline__ double eval_energy(const double *__restrict__ epart, int e, int ep_stride,
                                              int ep_n, double eself)
{
    return wave_sum_array(epart + (long)e * ep_stride, ep_n) - eself;
}

// Coefficients of the real-space special functions: the erfcx rational and
// the e^r Taylor series (e[k] = 1/k!).  As literals (LitCoef) they sit in
// VGPRs across the pair loop; passed as a kernel argument (RsCoef) they are
// loaded into SGPRs, which frees about 40 VGPRs (more waves per SIMD).
struct RsCoef {
    double p[QTP_ERFCX_DEG + 1], q[QTP_ERFCX_DEG + 1], e[14];
};
struct LitCoef {
    static constexpr double p[QTP_ERFCX_DEG + 1] = {QTP_ERFCX_P};
    static constexpr double q[QTP_ERFCX_DEG + 1] = {QTP_ERFCX_Q};
    static constexpr double e[14] = {1.0, 1.0, 0.5, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0,
                                     1.0 / 720.0, 1.0 / 5040.0, 1.0 / 40320.0, 1.0 / 362880.0,
                                     1.0 / 3628800.0, 1.0 / 39916800.0, 1.0 / 479001600.0,
                                     1.0 / 6227020800.0};
};
// the argument form, filled on the host with the same values
inline RsCoef rs_coef()
{
    RsCoef c;
    for (int k = 0; k <= QTP_ERFCX_DEG; ++k) {
        c.p[k] = LitCoef::p[k];
        c.q[k] = LitCoef::q[k];
    }
    for (int k = 0; k < 14; ++k) c.e[k] = LitCoef::e[k];
    return c;
}

// erfc(x) = exp(-x^2) erfcx(x) with erfcx from the P8/Q8 rational of
// qtp_erfcx.h (relative error < 9e-16 on [0, 6.2]; the host guarantees
// kappa * r_c <= 6.2, so every in-cutoff pair is in range).  ex = exp(-x^2)
// is passed in (the force needs it too).  The quotient uses the hardware
// reciprocal refined by two Newton steps and one residual correction
// (Q >= 1 on the range: no scaling cases).
template <class C>
__device__ __forceinline__ double erfc_via_erfcx(double x, double ex, const C &c)
{
    double p = c.p[QTP_ERFCX_DEG], q = c.q[QTP_ERFCX_DEG];
#pragma unroll
    for (int k = QTP_ERFCX_DEG - 1; k >= 0; --k) {
        p = fma(p, x, c.p[k]);
        q = fma(q, x, c.q[k]);
    }
    double r = __builtin_amdgcn_rcp(q);
    r = fma(fma(-q, r, 1.0), r, r);
    r = fma(fma(-q, r, 1.0), r, r);
    double y = p * r;
    y = fma(fma(-q, y, p), r, y);
    return ex * y;
}

// 1/sqrt(x) for finite x > 0: the hardware estimate (relative error below
// 2^-22) and one third-order Newton correction y (1 + e/2 + 3e^2/8) with
// e = 1 - x y^2, which leaves about one ulp.
__device__ __forceinline__ double rsqrt_pos(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-(x * y), y, 1.0);
    return fma(y * e, fma(0.375, e, 0.5), y);
}

// e^x for -708 <= x <= 0 (no overflow, no subnormal results in the range the
// real-space kernel uses, x >= -(kappa r_c)^2 >= -38.5): Cody-Waite reduction
// x = n ln2 + r, |r| <= ln2/2, with ln2 split as hi (its first 32 significant
// bits, so n hi is exact) + lo; e^r by its Taylor series to degree 13
// (truncation below 5e-18); then 2^n.  Under one ulp over [-40, 0].
template <class C>
__device__ __forceinline__ double exp_nonpos(double x, const C &c)
{
    constexpr double L2E = 1.4426950408889634;          // log2(e)
    constexpr double LN2_HI = 0.6931471803691238;       // ln 2 to 32 bits
    constexpr double LN2_LO = 1.9082149292705877e-10;   // ln 2 - LN2_HI
    const double n = rint(x * L2E);
    double r = fma(-n, LN2_HI, x);
    r = fma(-n, LN2_LO, r);
    double p = c.e[13];
#pragma unroll
    for (int k = 12; k >= 0; --k) p = fma(p, r, c.e[k]);
    return ldexp(p, (int)n);
}

}  // namespace qtp
