// qtp_kernels.hpp -- gfx950 device code of the qTIP4P/F path-integral step.
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

// Real-space Ewald over all site pairs i<j within r_c (forces.cpp:586-608,
// rwald :141-177).  One 256-thread block per 64x64 tile (I<=J) of the site
// pair matrix; wave w takes the 16 j-columns 16w..16w+15 against all 64 i.
// Pass 1 (lane <-> i) marks each lane's in-cutoff columns (16-bit mask).
// The wave then lays its in-cutoff pairs out in an LDS queue, lane-major
// (offsets by a wave prefix sum of the mask popcounts), and pass 2 gives
// every lane the same number of consecutive queue entries: lanes stay busy
// however unevenly the pairs fall on the rows.  Consecutive entries mostly
// share i, so a lane keeps its i-force in registers and adds it to the
// wave's LDS row when i changes; j reactions go to the LDS column array with
// ds_add_f64.  All LDS adds of a row or column come from one wavefront, in
// program order and fixed lane order, so the result is reproducible.
// Partial site forces go to slot J (block I's sites) and slot I (block J's
// sites): every (site, slot) has exactly one writer.
// grid: E * ntile, 256 threads.
// VIR: also accumulate the pair virial r_ij (x) f_ij (6 components) into
// vpart[(e*6 + c)*ep_stride + ep_off + tile].
// CARG: the special functions' coefficients come from the argument cf (SGPRs)
// instead of literals (VGPRs); the host picks it for large grids.
constexpr int RS_WAVES = 4;
constexpr long RS_CARG_BLOCKS = 16384;  // grids at least this large use CARG
template <bool VIR, bool CARG>
__global__ __launch_bounds__(256) void k_realspace(const double *__restrict__ pos, long pstride,
                                                   const double *__restrict__ charge, int n_atoms,
                                                   int NAP, int nblk, double gamma, double half1mg,
                                                   const int2 *__restrict__ tiles, int ntile,
                                                   Box box, Ewald ew,
                                                   double *__restrict__ rpart,
                                                   double *__restrict__ epart, int ep_stride,
                                                   int ep_off, double *__restrict__ vpart,
                                                   const RsCoef cf)
{
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int e = bid / ntile, t = bid - e * ntile;
    const int I = tiles[t].x, J = tiles[t].y;
    const int l = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const double *X = pos + (long)e * pstride;
    __shared__ double4 sI[64], sJ[64];  // x, y, z, charge
    __shared__ double fj[3][64];
    __shared__ double fi[RS_WAVES][3][64];
    __shared__ unsigned short qe[RS_WAVES][64 * 16];
    __shared__ double ewv[RS_WAVES];
    __shared__ double wv[RS_WAVES][6];
    if (w < 2) {
        const int s = (w == 0 ? J : I) * 64 + l;
        double4 v = make_double4(0.0, 0.0, 0.0, 0.0);
        if (s < n_atoms)
            v = make_double4(site_coord(X, NAP, s, 0, gamma, half1mg),
                             site_coord(X, NAP, s, 1, gamma, half1mg),
                             site_coord(X, NAP, s, 2, gamma, half1mg), charge[s]);
        (w == 0 ? sJ : sI)[l] = v;
        if (w == 0) {
            fj[0][l] = 0.0;
            fj[1][l] = 0.0;
            fj[2][l] = 0.0;
        }
    }
    fi[w][0][l] = 0.0;
    fi[w][1][l] = 0.0;
    fi[w][2][l] = 0.0;
    __syncthreads();
    const int c0 = 16 * w;
    const int i = I * 64 + l;
    // pass 1: in-cutoff columns of row i (j > i; j < n_atoms)
    unsigned int mask = 0u;
    if (i < n_atoms) {
        const double4 a = sI[l];
        for (int c = 0; c < 16; ++c) {
            const int jl = c0 + c;
            const int j = J * 64 + jl;
            const double4 b = sJ[jl];
            const double dx = mimg_fast(a.x - b.x, box.L[0], box.inv[0]);
            const double dy = mimg_fast(a.y - b.y, box.L[1], box.inv[1]);
            const double dz = mimg_fast(a.z - b.z, box.L[2], box.inv[2]);
            const double rsq = fma(dx, dx, fma(dy, dy, dz * dz));
            if (j > i && j < n_atoms && rsq < ew.rcut_sq) mask |= 1u << c;
        }
    }
    // the wave's pair queue, lane-major
    const int pc = __popc(mask);
    int incl = pc;
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (l >= o) incl += v;
    }
    const int total = __shfl(incl, 63, 64);
    {
        int pos_q = incl - pc;
        unsigned int m = mask;
        while (m) {
            const int c = __ffs(m) - 1;
            m &= m - 1u;
            qe[w][pos_q++] = (unsigned short)(l | (c << 6));
        }
    }
    __syncthreads();
    // pass 2: lane l takes queue entries [l*S, l*S + S)
    const bool near = J - I <= 1;
    const int S = (total + 63) >> 6;
    const int q0 = min(l * S, total), q1 = min(q0 + S, total);
    double fx = 0.0, fy = 0.0, fz = 0.0, en = 0.0;
    double vxx = 0.0, vyy = 0.0, vzz = 0.0, vxy = 0.0, vxz = 0.0, vyz = 0.0;
    int icur = -1;
    double4 a = make_double4(0.0, 0.0, 0.0, 0.0);
    for (int q = q0; q < q1; ++q) {
        const int ent = qe[w][q];
        const int il = ent & 63, jl = c0 + (ent >> 6);
        if (il != icur) {
            if (icur >= 0) {
                atomicAdd(&fi[w][0][icur], fx);
                atomicAdd(&fi[w][1][icur], fy);
                atomicAdd(&fi[w][2][icur], fz);
            }
            fx = fy = fz = 0.0;
            icur = il;
            a = sI[il];
        }
        const double4 b = sJ[jl];
        const double dx = mimg_fast(a.x - b.x, box.L[0], box.inv[0]);
        const double dy = mimg_fast(a.y - b.y, box.L[1], box.inv[1]);
        const double dz = mimg_fast(a.z - b.z, box.L[2], box.inv[2]);
        const double rsq = fma(dx, dx, fma(dy, dy, dz * dz));
        const double rinv = rsqrt_pos(rsq);
        double ex, ec;
        if constexpr (CARG) {
            ex = exp_nonpos(-ew.kappa2 * rsq, cf);
            ec = erfc_via_erfcx(ew.kappa * (rsq * rinv), ex, cf);
        } else {
            constexpr LitCoef lc{};
            ex = exp_nonpos(-ew.kappa2 * rsq, lc);
            ec = erfc_via_erfcx(ew.kappa * (rsq * rinv), ex, lc);
        }
        // same molecule: -erf (forces.cpp:158-166); a molecule's three sites
        // lie in one 64-site block or two consecutive ones
        if (near && (I * 64 + il) / 3 == (J * 64 + jl) / 3) ec -= 1.0;
        const double qq = a.w * b.w;
        const double ep = qq * rinv * ec;
        const double pf = (ep + qq * ew.pref2k * ex) * (rinv * rinv);
        en += ep;
        fx = fma(pf, dx, fx);
        fy = fma(pf, dy, fy);
        fz = fma(pf, dz, fz);
        if (VIR) {
            vxx = fma(pf * dx, dx, vxx);
            vyy = fma(pf * dy, dy, vyy);
            vzz = fma(pf * dz, dz, vzz);
            vxy = fma(pf * dx, dy, vxy);
            vxz = fma(pf * dx, dz, vxz);
            vyz = fma(pf * dy, dz, vyz);
        }
        atomicAdd(&fj[0][jl], -pf * dx);
        atomicAdd(&fj[1][jl], -pf * dy);
        atomicAdd(&fj[2][jl], -pf * dz);
    }
    if (icur >= 0) {
        atomicAdd(&fi[w][0][icur], fx);
        atomicAdd(&fi[w][1][icur], fy);
        atomicAdd(&fi[w][2][icur], fz);
    }
    en = wave_sum(en);
    if (l == 0) ewv[w] = en;
    if (VIR) {
        const double v6[6] = {vxx, vyy, vzz, vxy, vxz, vyz};
        for (int cc = 0; cc < 6; ++cc) {
            const double v = wave_sum(v6[cc]);
            if (l == 0) wv[w][cc] = v;
        }
    }
    __syncthreads();
    if (VIR && w == 0 && l < 6)
        vpart[((long)e * 6 + l) * ep_stride + ep_off + t] = wv[0][l] + wv[1][l] + wv[2][l] + wv[3][l];
    const long base_e = (long)e * nblk * 3 * NAP;
    if (w < 3) {
        // wave d writes component d of both slot entries of this tile
        const int d = w;
        const double si = fi[0][d][l] + fi[1][d][l] + fi[2][d][l] + fi[3][d][l];
        if (I == J) {
            rpart[base_e + (long)I * 3 * NAP + d * NAP + i] = si + fj[d][l];
        } else {
            rpart[base_e + (long)J * 3 * NAP + d * NAP + i] = si;                 // slot J
            rpart[base_e + (long)I * 3 * NAP + d * NAP + J * 64 + l] = fj[d][l];  // slot I
        }
    } else if (l == 0) {
        epart[(long)e * ep_stride + ep_off + t] = ewv[0] + ewv[1] + ewv[2] + ewv[3];
    }
}

// accumulator of v_mfma_f64_16x16x4f64: D[i][j] sits in lane j + 16 (i % 4),
// element i / 4; operands A[i][k] from lane i + 16 k, B[k][j] from lane j + 16 k
typedef double mf_v4d __attribute__((ext_vector_type(4)));

// Per-site factor table e^{i m b_d r_d}, m = 0..K-1, by the cos/sin recurrence
// of forces.cpp:212-245 (charge folded into the x factor, :241-244).
template <int STRIDE>
__device__ __forceinline__ void kw_table(double2 *row, double r, double bk, double qf, int K)
{
    const double kr = bk * r;
    double sn, cs;
    sincos(kr, &sn, &cs);
    double cm = 1.0, sm = 0.0;
    row[0] = make_double2(cm * qf, sm * qf);
    for (int m = 1; m < K; ++m) {
        const double cn = cs * cm - sn * sm;
        const double snn = sn * cm + cs * sm;
        cm = cn;
        sm = snn;
        row[m * STRIDE] = make_double2(cm * qf, sm * qf);
    }
}

constexpr int KF_REC = 9;   // doubles per k-group record read by k_kforce

// Reciprocal-space structure factors (kwald, forces.cpp:180-448), partial
// over a chunk of AC sites.  A k-"group" (kx>=0, ky>=0, kz>=0) carries the
// four sign variants (kx,+-ky,+-kz) of the reference's branches; lane <-> group.
// Output: 8 sums per group: Re/Im of A Z, A Z*, B Z, B Z* with A = X Y,
// B = X Y* (X,Y,Z the x,y,z factors).   grid: E * ngc * nac.
__global__ __launch_bounds__(256) void k_ksum(const double *__restrict__ pos, long pstride,
                                              const double *__restrict__ charge, int n_atoms,
                                              int NAP, double gamma, double half1mg, Ewald ew,
                                              const int4 *__restrict__ groups, int G, int AC,
                                              int ngc, int nac, double *__restrict__ spart)
{
    extern __shared__ double2 T[];  // [AC][3][K]
    const int per_e = ngc * nac;
    const int e = blockIdx.x / per_e;
    const int r = blockIdx.x - e * per_e;
    const int gc = r / nac, ac = r - gc * nac;
    const int K = ew.K;
    const double *X = pos + (long)e * pstride;
    for (int t = threadIdx.x; t < AC * 3; t += blockDim.x) {
        const int a = t / 3, d = t - 3 * a;
        const int atom = ac * AC + a;
        double2 *row = T + (a * 3 + d) * K;
        if (atom < n_atoms) {
            kw_table<1>(row, site_coord(X, NAP, atom, d, gamma, half1mg), ew.bk[d],
                        d == 0 ? charge[atom] : 1.0, K);
        } else {
            for (int m = 0; m < K; ++m) row[m] = make_double2(0.0, 0.0);
        }
    }
    __syncthreads();
    const int g = gc * blockDim.x + threadIdx.x;
    if (g >= G) return;
    const int4 gr = groups[g];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, s5 = 0.0, s6 = 0.0, s7 = 0.0;
#pragma unroll 4
    for (int a = 0; a < AC; ++a) {
        const double2 Xv = T[(a * 3 + 0) * K + gr.x];
        const double2 Yv = T[(a * 3 + 1) * K + gr.y];
        const double2 Zv = T[(a * 3 + 2) * K + gr.z];
        const double cc = Xv.x * Yv.x, csn = Xv.x * Yv.y, sc = Xv.y * Yv.x, ss = Xv.y * Yv.y;
        const double Ar = cc - ss, Ai = csn + sc, Br = cc + ss, Bi = sc - csn;
        s0 = fma(Ar, Zv.x, fma(-Ai, Zv.y, s0));
        s1 = fma(Ai, Zv.x, fma(Ar, Zv.y, s1));
        s2 = fma(Ar, Zv.x, fma(Ai, Zv.y, s2));
        s3 = fma(Ai, Zv.x, fma(-Ar, Zv.y, s3));
        s4 = fma(Br, Zv.x, fma(-Bi, Zv.y, s4));
        s5 = fma(Bi, Zv.x, fma(Br, Zv.y, s5));
        s6 = fma(Br, Zv.x, fma(Bi, Zv.y, s6));
        s7 = fma(Bi, Zv.x, fma(-Br, Zv.y, s7));
    }
    double *o = spart + (((long)e * nac + ac) * G + g) * 8;
    o[0] = s0; o[1] = s1; o[2] = s2; o[3] = s3;
    o[4] = s4; o[5] = s5; o[6] = s6; o[7] = s7;
}

// The partial structure factors of k_ksum on FP64 matrix cores.  Per chunk
// of AC sites and per column (kx, ky), the products
//   Q[4 col + p][n] = sum_j F_p(j) Zt[n][j],  F = (Xr Yr, Xi Yi, Xr Yi, Xi Yr)
//   at site j,  Zt[2 kz], Zt[2 kz + 1] = Re, Im Z(kz)
// form one matrix product over the sites; k_sreduce combines them into the
// eight sums of each group (A = X Y, B = X Y*: Ar = F0 - F1, Ai = F2 + F3,
// Br = F0 + F1, Bi = F3 - F2).  A unit = 16 rows (four columns in
// k_kforce_mf's sorted order) x 16 values of n (n-tile 0 or 1); each lane
// forms its A operand F_p with one product of two table entries.  Output per
// group and chunk (same place as k_ksum's sums): Re/Im Z times F0..F3, in
// that order.
// kcol[c] = (kx | ky << 16, first group, first kz, last kz) per sorted
// column (padding columns: first kz > last kz).  usplit blocks share a chunk
// (each builds its tables and takes 1/usplit of the units) when a launch has
// few chunks.
// grid: E * nac * usplit, 64 * KS_WAVES threads, ksum_mf_lds_bytes(K, AC, ncol4, nunit).
constexpr int KS_WAVES = 8;
// LDS rows of AC + 2 entries: the B operands of a quarter-wave (16 rows n at
// one site) and the A factors of a tile's columns then fall in different banks.
// The column and unit tables are staged in LDS behind the factor tables.
inline size_t ksum_mf_lds_bytes(int K, int AC, int ncol4, int nunit)
{
    return ((size_t)4 * K + 32) * (AC + 2) * 8 + (size_t)ncol4 * 16 + (size_t)nunit * 8;
}

__global__ __launch_bounds__(64 * KS_WAVES) void k_ksum_mf(
    const double *__restrict__ pos, long pstride, const double *__restrict__ charge, int n_atoms,
    int NAP, double gamma, double half1mg, Ewald ew, const int4 *__restrict__ kcol, int ncol4,
    const int2 *__restrict__ kunit, int nunit, int G, int AC, int nac, int usplit,
    double *__restrict__ spart)
{
    extern __shared__ double smem[];
    const int K = ew.K;
    const int RS = AC + 2;           // LDS row stride
    double2 *T = (double2 *)smem;    // [2][K][RS] x (charge folded), y factors
    double *Zt = smem + 4 * K * RS;  // [32][RS]
    int4 *sc = (int4 *)(Zt + 32 * RS);  // [ncol4] column table
    int2 *su = (int2 *)(sc + ncol4);    // [nunit] unit table
    const int per_e = nac * usplit;
    const int e = blockIdx.x / per_e, r = blockIdx.x - e * per_e;
    const int ac = r / usplit, us = r - ac * usplit;
    const int u0 = us * nunit / usplit, u1 = (us + 1) * nunit / usplit;
    const double *X = pos + (long)e * pstride;
    for (int t = threadIdx.x; t < ncol4; t += blockDim.x) sc[t] = kcol[t];
    for (int t = threadIdx.x; t < nunit; t += blockDim.x) su[t] = kunit[t];
    for (int t = threadIdx.x; t < 3 * AC; t += blockDim.x) {
        // factor tables by the recurrence of kw_table (forces.cpp:212-245)
        const int d = t / AC, j = t - d * AC;
        const int atom = ac * AC + j;
        const bool va = atom < n_atoms;
        double sn = 0.0, cs = 0.0, qf = 0.0;
        if (va) {
            qf = d == 0 ? charge[atom] : 1.0;
            sincos(ew.bk[d] * site_coord(X, NAP, atom, d, gamma, half1mg), &sn, &cs);
        }
        double cm = 1.0, sm = 0.0;
        for (int m = 0; m < K; ++m) {
            if (m > 0) {
                const double cn = cs * cm - sn * sm;
                const double snn = sn * cm + cs * sm;
                cm = cn;
                sm = snn;
            }
            const double c = va ? cm * qf : 0.0, s = va ? sm * qf : 0.0;
            if (d < 2) {
                T[(d * K + m) * RS + j] = make_double2(c, s);
            } else {
                Zt[(2 * m) * RS + j] = c;
                Zt[(2 * m + 1) * RS + j] = s;
            }
        }
        if (d == 2)
            for (int n = 2 * K; n < 32; ++n) Zt[n * RS + j] = 0.0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int j16 = lane & 15, gq = lane >> 4;
    // A row j16 = (column j16 / 4, product p = lane % 4): F_p = X_a Y_b with
    // a = Re for p = 0, 2, b = Re for p = 0, 3 (Im otherwise)
    const int pp = lane & 3;
    const int xa = (pp == 1 || pp == 3) ? 1 : 0, yb = (pp == 1 || pp == 2) ? 1 : 0;
    const double *Td = (const double *)T;
    double *out = spart + ((long)e * nac + ac) * G * 8;
    for (int u = u0 + w; u < u1; u += KS_WAVES) {
        const int2 un = su[u];
        const int ck = sc[4 * un.x + (j16 >> 2)].x;
        const int kx = ck & 0xffff, ky = ck >> 16;
        const double *Xr = Td + 2 * (kx * RS + gq) + xa;  // site 4 s + gq
        const double *Yr = Td + 2 * ((K + ky) * RS + gq) + yb;
        const double *Zr = Zt + (16 * un.y + j16) * RS + gq;
        mf_v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
        for (int s4 = 0; s4 < AC; s4 += 4)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Xr[2 * s4] * Yr[2 * s4], Zr[s4], acc, 0, 0, 0);
        // lane: n = 16 n-tile + j16, product gq, column 4 tile + reg
        const int n = 16 * un.y + j16, kz = n >> 1;
#pragma unroll
        for (int reg = 0; reg < 4; ++reg) {
            const int4 cr = sc[4 * un.x + reg];
            if (kz >= cr.z && kz <= cr.w)
                out[(long)(cr.y + kz - cr.z) * 8 + gq * 2 + (n & 1)] = acc[reg];
        }
    }
}

// Sum the partial structure factors, energy per group (forces.cpp:274, 315,
// 364, 428) and the force record of k_kforce from S' = 2 c S of the variants
// the reference evaluates (gr.w bit v set; others zero).
// Eight lanes per group, lane v <-> component v of the 8 partial sums, so the
// chunk partials are read as contiguous 64-byte rows; the chunks are summed
// in ascending order.  grid: E * ceil(G / SR_GPB), 256 threads.
// With vpart set, also the reciprocal-space virial tensor
//   W_ab = sum_k E_k (delta_ab - 2 k_a k_b (1/k^2 + 1/(4 kappa^2)))
// (E_k = c |S(k)|^2 per sign variant; the strain derivative of kwald's sum at
// fixed Ewald parameters).
// With wm set, the force coefficients go to k_kforce_mf's matrix instead of
// the records: W[n][8 col + s] (n = 2 kz + 0/1 for the Re/Im z factor; s =
// u1 u2 v1 v2 u3 u4 v3 v4, the Z-linear sums of k_kforce; col = gcol[g]).
// Each coefficient c_v enters twice: as the u weight of its own component and,
// times +-kz b_z, as the v weight of the other part of the same variant pair.
// Only these entries are ever written; the rest of W stays zero.
constexpr int SR_GPB = 32;  // groups per block
__global__ __launch_bounds__(256) void k_sreduce(const double *__restrict__ spart, int nac, int G,
                                                 const int4 *__restrict__ groups,
                                                 const double *__restrict__ gcf,
                                                 double *__restrict__ sprime,
                                                 double *__restrict__ epart, int ep_stride,
                                                 int ep_off, const double4 *__restrict__ grk,
                                                 double q4, double *__restrict__ vpart,
                                                 double *__restrict__ wm, int Mp, int NN,
                                                 const int *__restrict__ gcol, double bkz,
                                                 int pform)
{
    __shared__ double sh[16];
    const int nb = (G + SR_GPB - 1) / SR_GPB;
    const int e = blockIdx.x / nb, gb = blockIdx.x - e * nb;
    const int g = gb * SR_GPB + (threadIdx.x >> 3);
    const int v = threadIdx.x & 7;
    const bool live = g < G;
    double sv = 0.0, cf = 0.0;
    if (live) {
        const double *p = spart + ((long)e * nac * G + g) * 8 + v;
        const long cs = (long)G * 8;
        int ac = 0;
        // eight chunk loads in flight, summed in ascending chunk order
        for (; ac + 8 <= nac; ac += 8) {
            double t[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) t[j] = p[(ac + j) * cs];
#pragma unroll
            for (int j = 0; j < 8; ++j) sv += t[j];
        }
        for (; ac < nac; ++ac) sv += p[ac * cs];
    }
    if (pform) {
        // k_ksum_mf's sums q = (Re Z, Im Z) x (Xr Yr, Xi Yi, Xr Yi, Xi Yr) ->
        // the sums of k_ksum: Re/Im of A Z, A Z*, then B Z, B Z*
        const int base = threadIdx.x & 63 & ~7;
        double q[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) q[i] = __shfl(sv, base + i, 64);
        const double ArZr = q[0] - q[2], AiZi = q[5] + q[7], AiZr = q[4] + q[6], ArZi = q[1] - q[3];
        const double BrZr = q[0] + q[2], BiZi = q[7] - q[5], BiZr = q[6] - q[4], BrZi = q[1] + q[3];
        switch (v) {
        case 0: sv = ArZr - AiZi; break;
        case 1: sv = AiZr + ArZi; break;
        case 2: sv = ArZr + AiZi; break;
        case 3: sv = AiZr - ArZi; break;
        case 4: sv = BrZr - BiZi; break;
        case 5: sv = BiZr + BrZi; break;
        case 6: sv = BrZr + BiZi; break;
        default: sv = BiZr - BrZi; break;
        }
    }
    if (live) {
        if (!((groups[g].w >> (v >> 1)) & 1)) sv = 0.0;
        cf = gcf[g];
    }
    double en = cf * sv * sv;
    double v6[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    // the four values of this lane's side: S0r, S0i, S1r, S1i (times 2c)
    const int base = (threadIdx.x & 63) & ~3;
    const double c2 = 2.0 * cf;
    const double S0r = c2 * __shfl(sv, base + 0, 64), S0i = c2 * __shfl(sv, base + 1, 64);
    const double S1r = c2 * __shfl(sv, base + 2, 64), S1i = c2 * __shfl(sv, base + 3, 64);
    if (vpart) {
        const double pair = __shfl_xor(sv, 1, 64);
        if (live && !(v & 1)) {
            const double4 rk = grk[g];
            const double k2 = rk.x * rk.x + rk.y * rk.y + rk.z * rk.z;
            const double bb = 2.0 * (1.0 / k2 + q4);
            const int var = v >> 1;
            const double Ev = cf * (sv * sv + pair * pair);
            const double ky = (var < 2) ? rk.y : -rk.y, kz = (var & 1) ? -rk.z : rk.z;
            v6[0] = Ev * (1.0 - bb * rk.x * rk.x);
            v6[1] = Ev * (1.0 - bb * ky * ky);
            v6[2] = Ev * (1.0 - bb * kz * kz);
            v6[3] = -Ev * bb * rk.x * ky;
            v6[4] = -Ev * bb * rk.x * kz;
            v6[5] = -Ev * bb * ky * kz;
        }
    }
    if (live) {
        // force coefficients: A side (variants 0, 1): c1 = S0i+S1i,
        // c2 = S1r-S0r, c3 = -(S0r+S1r), c4 = S1i-S0i; B side (variants 2, 3)
        // likewise
        const int k = v & 3;
        const double cv = k == 0 ? S0i + S1i : k == 1 ? S1r - S0r : k == 2 ? -(S0r + S1r) : S1i - S0i;
        if (wm) {
            const int kz = groups[g].z;
            const int lo = v & 1, hi = (v >> 1) & 1;
            const int row = 8 * gcol[g] + (v >> 2) * 4;
            const double kzr = bkz * kz;
            double *W = wm + (long)e * NN * Mp;
            W[(long)(2 * kz + lo) * Mp + row + hi] = cv;
            W[(long)(2 * kz + 1 - lo) * Mp + row + 3 - hi] = (lo == hi ? -kzr : kzr) * cv;
        } else {
            // record of k_kforce (9 doubles): c1..c8, then (kx, ky, kz) packed
            double *o = sprime + ((long)e * G + g) * KF_REC;
            o[v] = cv;
            if (v == 0) {
                const int4 gg = groups[g];
                o[8] = __longlong_as_double((long long)(gg.x | (gg.y << 10) | (gg.z << 20)));
            }
        }
    }
    en = block_sum(en, sh);
    if (threadIdx.x == 0) epart[(long)e * ep_stride + ep_off + gb] = en;
    if (vpart) {
        for (int c = 0; c < 6; ++c) {
            const double vs = block_sum(v6[c], sh);
            if (threadIdx.x == 0) vpart[((long)e * 6 + c) * ep_stride + ep_off + gb] = vs;
        }
    }
}

// Reciprocal-space forces (forces.cpp:276-279, 319-324, 368-373, 434-442),
// plus the real-space partial slots, the M-site chain rule (forces.cpp:613-619)
// and the LJ+intra forces: the physical force of every site of a block of
// 21 molecules (63 sites, lane <-> site; lane 63 idle), the waves splitting
// the k-groups (details at k_kforce).
constexpr int KF_WAVES = 12;   // waves per block: two blocks fit one CU
// Split form (the resident SC phases): each 63-site block is two blocks of
// KF_SPLIT_WAVES waves, one per half of the k-groups; the second half's
// force goes to a second buffer that the SC kernels add.  Three such blocks
// fit one CU, so the grid spreads over the CUs more evenly.
constexpr int KF_SPLIT_WAVES = 6;

constexpr int KF_SITES = 63;
constexpr int KF_CHUNK = 48;   // most k-group records staged per wave and round
// Records staged per wave and round for W-wave blocks and K-entry tables: as
// many as keep `per_cu` blocks within one CU's 160 KB of LDS, at most
// KF_CHUNK (large k_max leaves less room next to the factor tables); the
// area also holds the W x 3 x 64 reduction, which sets its floor.
inline int kforce_chunk(int K, int W, int per_cu)
{
    const long room = 160L * 1024 / per_cu - 3L * K * 64 * 16 - 3L * 64 * 8;
    const long c = std::max(room, (long)W * 3 * 64 * 8) / ((long)W * KF_REC * 8);
    return (int)std::max(8L, std::min((long)KF_CHUNK, c));
}
// LDS bytes of k_kforce for a K-entry factor table
inline size_t kforce_lds_bytes(int K, int W = KF_WAVES, int chunk = KF_CHUNK)
{
    const size_t cf = (size_t)W * chunk * KF_REC;
    const size_t red = (size_t)W * 3 * 64;
    return (size_t)3 * K * 64 * 16 + (cf > red ? cf : red) * 8 + (size_t)3 * 64 * 8;
}
// Reciprocal-space forces and the final per-atom forces of E evaluations.
// Block = 63 sites (21 whole molecules) x W waves; lanes <-> sites, wave w
// takes a contiguous part of the k-groups.  Groups in build order
// form columns of fixed (kx, ky) and rising kz.  For site j the group's
// force is minus the k-weighted imaginary parts of S'_v* e^{i k_v r_j} over
// its sign variants; with A = X Y, B = X Y* (per column) and Z the z factor
// it is linear in Z:
//   p01 = Ar (c1 Zr + c2 Zi) + Ai (c3 Zr + c4 Zi)          (fx, fy; variants 0, 1)
//   m01 = Ar kz (-c4 Zr + c3 Zi) + Ai kz (c2 Zr - c1 Zi)   (fz)
// and p23, m23 likewise with B and c5..c8.  A column accumulates the four
// Z-linear sums per side over its groups and applies A, B, kx, ky once.
// Group records (k_sreduce) are staged per wave in LDS with vector loads,
// chunk (<= KF_CHUNK) at a time; the real-space slot partials are loaded at kernel
// start (wave w takes slots w, w + KF_WAVES, ...).  Epilogue: waves 0-2 (one
// per dimension) sum the wave partials, apply the M-site chain rule
// (forces.cpp:613-619) and add the LJ/intramolecular forces (and the
// splined pair forces when fspl is given).
// grid: E * ceil(n_atoms / 63), 64 * KF_WAVES threads, kforce_lds_bytes(K).
// SPLIT: grid 2 * E * nab; blocks [0, E nab) take the first half of the
// groups and everything else, writing fout; blocks [E nab, 2 E nab) take the
// second half of the groups only, writing fout2.
template <int W, bool SPLIT>
__global__ __launch_bounds__(64 * W) void k_kforce(
    const double *__restrict__ pos, long pstride, const double *__restrict__ charge, int n_atoms,
    int NAP, double gamma, double half1mg, Ewald ew, int G, const double *__restrict__ krec,
    const double *__restrict__ rpart, int nblk, const double *__restrict__ flj,
    const double *__restrict__ fspl, double *__restrict__ fout, double *__restrict__ fout2,
    int chunk)
{
    constexpr int KF_WAVES = W;
    extern __shared__ double smem[];
    const int K = ew.K;
    double2 *T = (double2 *)smem;           // [3][K][64]
    double *CF = smem + 2 * 3 * K * 64;     // [KF_WAVES][chunk][KF_REC]; later red [W][3][64]
    const size_t cfn = (size_t)KF_WAVES * chunk * KF_REC;
    const size_t redn = (size_t)KF_WAVES * 3 * 64;
    double *ef = CF + (cfn > redn ? cfn : redn);  // [3][64]
    const int nab = (n_atoms + KF_SITES - 1) / KF_SITES;
    int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int half = SPLIT ? (bid >= (int)(gridDim.x >> 1) ? 1 : 0) : 0;
    if (SPLIT) bid -= half * (int)(gridDim.x >> 1);
    const int e = bid / nab, ab = bid - e * nab;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int atom = ab * KF_SITES + lane;
    const bool va = lane < KF_SITES && atom < n_atoms;
    const double *X = pos + (long)e * pstride;
    double fx = 0.0, fy = 0.0, fz = 0.0, fl = 0.0;
    if (va && half == 0) {
        for (int sl = w; sl < nblk; sl += KF_WAVES) {
            const double *rp = rpart + ((long)e * nblk + sl) * 3 * NAP + atom;
            fx += rp[0];
            fy += rp[NAP];
            fz += rp[2 * NAP];
        }
        if (w < 3) {
            fl = flj[((long)e * 3 + w) * NAP + atom];
            if (fspl) fl += fspl[((long)e * 3 + w) * NAP + atom];
        }
    }
    if (threadIdx.x < 192) {
        const int d = threadIdx.x >> 6;
        if (va)
            kw_table<64>(T + d * K * 64 + lane, site_coord(X, NAP, atom, d, gamma, half1mg),
                         ew.bk[d], d == 0 ? charge[atom] : 1.0, K);
        else
            for (int m = 0; m < K; ++m) T[(d * K + m) * 64 + lane] = make_double2(0.0, 0.0);
    }
    constexpr int VW = SPLIT ? 2 * KF_WAVES : KF_WAVES;  // group shares
    const int per = (G + VW - 1) / VW;
    const int g0 = (half * KF_WAVES + w) * per, g1 = min(G, g0 + per);
    const int rounds = (per + chunk - 1) / chunk;
    double *cw = CF + (size_t)w * chunk * KF_REC;
    const double *R = krec + (long)e * G * KF_REC;
    double Ar = 0.0, Ai = 0.0, Br = 0.0, Bi = 0.0, kxr = 0.0, kyr = 0.0;
    double u1 = 0.0, u2 = 0.0, v1 = 0.0, v2 = 0.0, u3 = 0.0, u4 = 0.0, v3 = 0.0, v4 = 0.0;
    int pkxy = -1;
    for (int rd = 0; rd < rounds; ++rd) {
        const int c0 = g0 + rd * chunk;
        const int nc = max(0, min(chunk, g1 - c0));
        for (int t = lane; t < nc * KF_REC; t += 64) cw[t] = R[(long)c0 * KF_REC + t];
        __syncthreads();  // tables (first round) and this round's records
        for (int i = 0; i < nc; ++i) {
            const double *c = cw + i * KF_REC;
            // the same for every lane: made scalar so the column logic is uniform
            const int tag = __builtin_amdgcn_readfirstlane((int)__double_as_longlong(c[8]));
            const int kx = tag & 1023, ky = (tag >> 10) & 1023, kz = tag >> 20;
            const int kxy = kx | (ky << 10);
            if (kxy != pkxy) {
                // close the previous column
                const double p01 = fma(Ar, u1, Ai * u2), m01 = fma(Ar, v1, Ai * v2);
                const double p23 = fma(Br, u3, Bi * u4), m23 = fma(Br, v3, Bi * v4);
                fx = fma(-kxr, p01 + p23, fx);
                fy = fma(-kyr, p01 - p23, fy);
                fz -= m01 + m23;
                u1 = u2 = v1 = v2 = u3 = u4 = v3 = v4 = 0.0;
                const double2 Xv = T[(0 * K + kx) * 64 + lane];
                const double2 Yv = T[(1 * K + ky) * 64 + lane];
                const double cc = Xv.x * Yv.x, csn = Xv.x * Yv.y, sc = Xv.y * Yv.x,
                             ss = Xv.y * Yv.y;
                Ar = cc - ss;
                Ai = csn + sc;
                Br = cc + ss;
                Bi = sc - csn;
                kxr = ew.bk[0] * kx;
                kyr = ew.bk[1] * ky;
                pkxy = kxy;
            }
            const double2 Zv = T[(2 * K + kz) * 64 + lane];
            const double kzr = ew.bk[2] * kz;
            const double zkr = kzr * Zv.x, zki = kzr * Zv.y;
            u1 = fma(c[0], Zv.x, fma(c[1], Zv.y, u1));
            u2 = fma(c[2], Zv.x, fma(c[3], Zv.y, u2));
            v1 = fma(-c[3], zkr, fma(c[2], zki, v1));
            v2 = fma(c[1], zkr, fma(-c[0], zki, v2));
            if (ky > 0) {  // B-side variants exist only off the ky = 0 plane
                u3 = fma(c[4], Zv.x, fma(c[5], Zv.y, u3));
                u4 = fma(c[6], Zv.x, fma(c[7], Zv.y, u4));
                v3 = fma(-c[7], zkr, fma(c[6], zki, v3));
                v4 = fma(c[5], zkr, fma(-c[4], zki, v4));
            }
        }
        __syncthreads();  // records consumed before the next round (or red) overwrites them
    }
    if (rounds == 0) __syncthreads();
    {
        const double p01 = fma(Ar, u1, Ai * u2), m01 = fma(Ar, v1, Ai * v2);
        const double p23 = fma(Br, u3, Bi * u4), m23 = fma(Br, v3, Bi * v4);
        fx = fma(-kxr, p01 + p23, fx);
        fy = fma(-kyr, p01 - p23, fy);
        fz -= m01 + m23;
    }
    double *red = CF;  // [KF_WAVES][3][64]
    red[(w * 3 + 0) * 64 + lane] = fx;
    red[(w * 3 + 1) * 64 + lane] = fy;
    red[(w * 3 + 2) * 64 + lane] = fz;
    __syncthreads();
    if (w < 3) {  // wave d: total Ewald site force along d
        double v = 0.0;
#pragma unroll
        for (int ww = 0; ww < KF_WAVES; ++ww) v += red[(ww * 3 + w) * 64 + lane];
        ef[w * 64 + lane] = v;
    }
    __syncthreads();
    if (w < 3 && va) {
        const int site = lane % 3;  // 63 = 21 molecules: lane%3 == atom%3
        const double v = site == 0 ? fl + gamma * ef[w * 64 + lane]
                                   : fl + (ef[w * 64 + lane] + half1mg * ef[w * 64 + lane - site]);
        (half ? fout2 : fout)[((long)e * 3 + w) * NAP + atom] = v;
    }
}

// Reciprocal-space forces on FP64 matrix cores (v_mfma_f64_16x16x4f64), the
// same sums as k_kforce (forces.cpp:276-279, 319-324, 368-373, 434-442).
// Per evaluation, the Z-linear sums of every column and every site form one
// product  D[8 col + s][site] = sum_n W[8 col + s][n] Zt[n][site]  with
// Zt[2 kz][j], Zt[2 kz + 1][j] = Re, Im e^{i kz b_z z_j} and W written by
// k_sreduce.  Columns are sorted by their largest kz, so a 16-row tile (two
// columns) takes only the k-steps of 4 that its columns reach (tile table
// rt: packed (kx, ky) of both columns, k-steps, first row; in the order the
// waves take them, k-steps balanced over the waves; k-steps 0 = no tile).
// The epilogue applies each column's A = XY, B = XY* and kx, ky per site:
// with lane group g = lane / 16 holding rows 8 c' + g and 8 c' + g + 4 of
// both columns of a tile (the 16x16 result layout), g = 0, 1 give the x and
// y forces of (Ar u1 + Br u3) and (Ai u2 + Bi u4), g = 2, 3 the z force of
// (Ar v1 + Br v3) and (Ai v2 + Bi v4).
// Block = 48 sites (16 whole molecules: three 16-site tiles) x W waves (8,
// or 16 when a launch has few site blocks); each wave reads the z factors of
// the three site tiles (the B operands) from LDS and takes every W-th row
// tile (the tile table is ordered for W).  Epilogue as k_kforce: wave
// partials summed in wave order, real-space slot partials, M-site chain rule
// (forces.cpp:613-619), LJ/intra (and spline) forces.
// grid: E * ceil(n_atoms / 48), 64 * W threads, kforce_mf_lds_bytes(K, nks, W).
constexpr int MF_SITES = 48;
constexpr int MF_WAVES = 8;       // waves per block of large launches
constexpr int MF_WAVES_SMALL = 16;  // and of launches of fewer than MF_SMALL_BLOCKS blocks
constexpr int MF_SMALL_BLOCKS = 64;
constexpr int MF_MAXKS = 8;  // k-steps per tile (A operands in registers): 2 K <= 32 (k_max <= 15)

inline size_t kforce_mf_lds_bytes(int K, int nks, int W)
{
    return (size_t)2 * K * MF_SITES * 16 + (size_t)4 * nks * MF_SITES * 8 +
           (size_t)2 * W * 3 * MF_SITES * 8 + (size_t)3 * MF_SITES * 8;
}

// One 16-row tile of k_kforce_mf over three site tiles, NT k-steps: the A
// operands (W rows, global), then the B operands (z factors, LDS), then the
// 3 NT products, as straight-line code so that the loads overlap.
template <int NT>
__device__ __forceinline__ void mf_tile(const double *__restrict__ Wr, long Mp4,
                                        const double *__restrict__ Zl, mf_v4d &d0, mf_v4d &d1,
                                        mf_v4d &d2)
{
    double a[NT], b0[NT], b1[NT], b2[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) a[t] = Wr[t * Mp4];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        b0[t] = Zl[4 * t * MF_SITES];
        b1[t] = Zl[4 * t * MF_SITES + 16];
        b2[t] = Zl[4 * t * MF_SITES + 32];
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        d0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b0[t], d0, 0, 0, 0);
        d1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b1[t], d1, 0, 0, 0);
        d2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b2[t], d2, 0, 0, 0);
    }
}

template <int W>
__global__ __launch_bounds__(64 * W) void k_kforce_mf(
    const double *__restrict__ pos, long pstride, const double *__restrict__ charge, int n_atoms,
    int NAP, double gamma, double half1mg, Ewald ew, const double *__restrict__ wm, int Mp,
    int nks, const int4 *__restrict__ rt, int nrt, const double *__restrict__ rpart, int nblk,
    const double *__restrict__ flj, const double *__restrict__ fspl, double *__restrict__ fout)
{
    extern __shared__ double smem[];
    const int K = ew.K, NN = 4 * nks;
    double2 *T = (double2 *)smem;                    // [2][K][48] x (charge folded), y factors
    double *Zt = smem + 4 * K * MF_SITES;            // [NN][48]
    double *red = Zt + NN * MF_SITES;                // [W][3][48] k-space wave partials
    double *rsm = red + W * 3 * MF_SITES;     // [W][3][48] real-space slot partials
    double *ef = rsm + W * 3 * MF_SITES;      // [3][48]
    const int nbm = (n_atoms + MF_SITES - 1) / MF_SITES;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int e = bid / nbm, ab = bid - e * nbm;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int a0 = ab * MF_SITES;
    const double *X = pos + (long)e * pstride;
    double rx = 0.0, ry = 0.0, rz = 0.0;
    if (lane < MF_SITES && a0 + lane < n_atoms) {
        for (int sl = w; sl < nblk; sl += W) {
            const double *rp = rpart + ((long)e * nblk + sl) * 3 * NAP + a0 + lane;
            rx += rp[0];
            ry += rp[NAP];
            rz += rp[2 * NAP];
        }
    }
    if (threadIdx.x < 3 * MF_SITES) {
        // factor tables by the recurrence of kw_table (forces.cpp:212-245)
        const int d = threadIdx.x / MF_SITES, j = threadIdx.x - d * MF_SITES;
        const int atom = a0 + j;
        const bool va = atom < n_atoms;
        double sn = 0.0, cs = 0.0, qf = 0.0;
        if (va) {
            qf = d == 0 ? charge[atom] : 1.0;
            sincos(ew.bk[d] * site_coord(X, NAP, atom, d, gamma, half1mg), &sn, &cs);
        }
        double cm = 1.0, sm = 0.0;
        for (int m = 0; m < K; ++m) {
            if (m > 0) {
                const double cn = cs * cm - sn * sm;
                const double snn = sn * cm + cs * sm;
                cm = cn;
                sm = snn;
            }
            const double c = va ? cm * qf : 0.0, s = va ? sm * qf : 0.0;
            if (d < 2) {
                T[(d * K + m) * MF_SITES + j] = make_double2(c, s);
            } else {
                Zt[(2 * m) * MF_SITES + j] = c;
                Zt[(2 * m + 1) * MF_SITES + j] = s;
            }
        }
        if (d == 2)
            for (int n = 2 * K; n < NN; ++n) Zt[n * MF_SITES + j] = 0.0;
    }
    __syncthreads();
    const int j = lane & 15, gq = lane >> 4;
    const double *Zl = Zt + gq * MF_SITES + j;  // B operand of k-step t, site tile st: Zl[4t*48 + 16st]
    const bool odd = gq & 1;
    const double sg = odd ? 1.0 : -1.0;
    const bool gxy = gq < 2;
    double acc1[3] = {0.0, 0.0, 0.0}, acc2[3] = {0.0, 0.0, 0.0};
    const double *We = wm + (long)e * NN * Mp;
    for (int i = w; i < nrt; i += W) {
        const int4 r = rt[i];
        if (r.z == 0) continue;
        mf_v4d d0 = {0.0, 0.0, 0.0, 0.0}, d1 = d0, d2 = d0;
        const double *Wr = We + (long)gq * Mp + r.w + j;
        const long Mp4 = 4L * Mp;
        switch (r.z) {  // 1 <= r.z <= MF_MAXKS (host table)
        case 1: mf_tile<1>(Wr, Mp4, Zl, d0, d1, d2); break;
        case 2: mf_tile<2>(Wr, Mp4, Zl, d0, d1, d2); break;
        case 3: mf_tile<3>(Wr, Mp4, Zl, d0, d1, d2); break;
        case 4: mf_tile<4>(Wr, Mp4, Zl, d0, d1, d2); break;
        case 5: mf_tile<5>(Wr, Mp4, Zl, d0, d1, d2); break;
        case 6: mf_tile<6>(Wr, Mp4, Zl, d0, d1, d2); break;
        case 7: mf_tile<7>(Wr, Mp4, Zl, d0, d1, d2); break;
        default: mf_tile<8>(Wr, Mp4, Zl, d0, d1, d2); break;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int ck = h ? r.y : r.x;
            const int kx = ck & 0xffff, ky = ck >> 16;
            const double cx = gxy ? ew.bk[0] * kx : 1.0, cy = gxy ? ew.bk[1] * ky : 0.0;
#pragma unroll
            for (int st = 0; st < 3; ++st) {
                const mf_v4d &D = st == 0 ? d0 : st == 1 ? d1 : d2;
                const double2 Xv = T[kx * MF_SITES + 16 * st + j];
                const double2 Yv = T[(K + ky) * MF_SITES + 16 * st + j];
                const double s1 = Xv.x * (odd ? Yv.y : Yv.x), s2 = Xv.y * (odd ? Yv.x : Yv.y);
                const double a = fma(sg, s2, s1), b = fma(-sg, s1, s2);
                const double ta = a * D[2 * h], tb = b * D[2 * h + 1];
                acc1[st] = fma(-cx, ta + tb, acc1[st]);
                acc2[st] = fma(-cy, ta - tb, acc2[st]);
            }
        }
    }
#pragma unroll
    for (int st = 0; st < 3; ++st) {
        acc1[st] += __shfl_xor(acc1[st], 16, 64);
        acc2[st] += __shfl_xor(acc2[st], 16, 64);
        if (gq == 0) {
            red[(w * 3 + 0) * MF_SITES + 16 * st + j] = acc1[st];
            red[(w * 3 + 1) * MF_SITES + 16 * st + j] = acc2[st];
        } else if (gq == 2) {
            red[(w * 3 + 2) * MF_SITES + 16 * st + j] = acc1[st];
        }
    }
    if (lane < MF_SITES) {
        rsm[(w * 3 + 0) * MF_SITES + lane] = rx;
        rsm[(w * 3 + 1) * MF_SITES + lane] = ry;
        rsm[(w * 3 + 2) * MF_SITES + lane] = rz;
    }
    __syncthreads();
    const int d = threadIdx.x / MF_SITES, sj = threadIdx.x - d * MF_SITES;
    if (threadIdx.x < 3 * MF_SITES) {
        double vr = 0.0, vk = 0.0;
#pragma unroll
        for (int ww = 0; ww < W; ++ww) {
            vr += rsm[(ww * 3 + d) * MF_SITES + sj];
            vk += red[(ww * 3 + d) * MF_SITES + sj];
        }
        ef[d * MF_SITES + sj] = vr + vk;
    }
    __syncthreads();
    if (threadIdx.x < 3 * MF_SITES && a0 + sj < n_atoms) {
        const int atom = a0 + sj;
        const long o = ((long)e * 3 + d) * NAP + atom;
        double fl = flj[o];
        if (fspl) fl += fspl[o];
        const int site = sj % 3;  // 48 sites = 16 molecules: sj%3 == atom%3
        const double v = site == 0 ? fl + gamma * ef[d * MF_SITES + sj]
                                   : fl + (ef[d * MF_SITES + sj] + half1mg * ef[d * MF_SITES + sj - site]);
        fout[o] = v;
    }
}

// LJ over the full O-O list (forces.cpp:489-515, lj_oxygen :25-59) and the
// intramolecular terms (forces.cpp:519-554, :65-138).  LPM lanes per molecule
// share its list row (16, or a whole wave when the launch is small: one
// evaluation of a few hundred molecules leaves most CUs idle otherwise, and
// each lane's walk of the row is a chain of dependent loads); lane 0 of the
// group adds the intramolecular terms.  Each LJ pair is met from both sides;
// its energy is halved.
// grid: E * ceil(n_mols / (256 / LPM)), 256 threads.
template <bool VIR, int LPM>
__global__ __launch_bounds__(256) void k_ljintra(const double *__restrict__ pos, long pstride,
                                                 int n_mols, int NAP,
                                                 const int *__restrict__ fpoint,
                                                 const int *__restrict__ flist, long fcap,
                                                 int lmode, int lbase, Box box, FF ff,
                                                 double *__restrict__ flj,
                                                 double *__restrict__ epart, int ep_stride,
                                                 int ep_off, double *__restrict__ vpart)
{
    __shared__ double sh[16];
    constexpr int MPB = 256 / LPM;
    double v6[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int nmb = (n_mols + MPB - 1) / MPB;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int e = bid / nmb, mb = bid - e * nmb;
    const int gi = threadIdx.x / LPM, t = threadIdx.x % LPM;
    const int i = mb * MPB + gi;
    const double *P = pos + (long)e * pstride;
    double fo[3] = {0.0, 0.0, 0.0}, elj = 0.0, en = 0.0;
    double o[3] = {0.0, 0.0, 0.0};
    const int a = 3 * i;
    if (i < n_mols) {
        const int b = lst_bead(lmode, lbase, e);
        for (int d = 0; d < 3; ++d) o[d] = P[d * NAP + a];
        const int *fp = fpoint + (long)b * (n_mols + 1);
        const int *fl = flist + (long)b * fcap;
        // an overflowed list (status flag set by the list kernels) is never read past capacity
        const int k0 = (int)min((long)fp[i], fcap), k1 = (int)min((long)fp[i + 1], fcap);
        for (int k = k0 + t; k < k1; k += LPM) {
            const int aj = 3 * fl[k];
            double r[3];
            for (int d = 0; d < 3; ++d)
                r[d] = mimg_fast(o[d] - P[d * NAP + aj], box.L[d], box.inv[d]);
            const double rsq = fma(r[0], r[0], fma(r[1], r[1], r[2] * r[2]));
            if (rsq < ff.r_cut_sq) {
                const double rm2 = 1.0 / rsq;
                const double rm6 = rm2 * rm2 * rm2;
                const double a12 = rm6 * rm6 * ff.sigmalj12;
                const double a6 = rm6 * ff.sigmalj6;
                const double u = a12 - a6;
                elj += u * ff.epslj4 - ff.e_cut;
                const double pf = (u + a12) * ff.epslj24 * rm2;
                for (int d = 0; d < 3; ++d) fo[d] = fma(pf, r[d], fo[d]);
                if (VIR) {  // each pair is met twice: half of r (x) f per side
                    const double hp = 0.5 * pf;
                    v6[0] += hp * r[0] * r[0];
                    v6[1] += hp * r[1] * r[1];
                    v6[2] += hp * r[2] * r[2];
                    v6[3] += hp * r[0] * r[1];
                    v6[4] += hp * r[0] * r[2];
                    v6[5] += hp * r[1] * r[2];
                }
            }
        }
    }
    for (int off = LPM / 2; off > 0; off >>= 1) {
        for (int d = 0; d < 3; ++d) fo[d] += __shfl_xor(fo[d], off, 64);
        elj += __shfl_xor(elj, off, 64);
    }
    if (i < n_mols && t == 0) {
        double h1[3], h2[3];
        for (int d = 0; d < 3; ++d) {
            h1[d] = P[d * NAP + a + 1];
            h2[d] = P[d * NAP + a + 2];
        }
        double fh1[3] = {0.0, 0.0, 0.0}, fh2[3] = {0.0, 0.0, 0.0};
        en = 0.5 * elj;
        // intramolecular (forces.cpp:519-554)
        double r1[3], r2[3];
        for (int d = 0; d < 3; ++d) {
            r1[d] = mimg_fast(o[d] - h1[d], box.L[d], box.inv[d]);
            r2[d] = mimg_fast(o[d] - h2[d], box.L[d], box.inv[d]);
        }
        const double roh1 = sqrt(r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2]);
        const double roh2 = sqrt(r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2]);
        // bonds (intra_morse_bond)
        for (int bnd = 0; bnd < 2; ++bnd) {
            const double roh = bnd ? roh2 : roh1;
            const double *rr = bnd ? r2 : r1;
            double *fh = bnd ? fh2 : fh1;
            const double rm1 = 1.0 / roh;
            const double x = roh - ff.roheq;
            const double x2 = x * x, x3 = x2 * x, x4 = x2 * x2;
            en += (ff.ar2 * x2 - ff.ar3 * x3 + ff.ar4fac * x4) * ff.Dr;
            const double pf = (ff.ar2x2 * x - ff.ar3x3 * x2 + ff.ar4x4 * x3) * (ff.Dr * rm1);
            for (int d = 0; d < 3; ++d) {
                fo[d] -= pf * rr[d];
                fh[d] += pf * rr[d];
            }
            if (VIR) {  // r_O (x) F_O + r_H (x) F_H = -pf r (x) r, r = r_O - r_H
                v6[0] -= pf * rr[0] * rr[0];
                v6[1] -= pf * rr[1] * rr[1];
                v6[2] -= pf * rr[2] * rr[2];
                v6[3] -= pf * rr[0] * rr[1];
                v6[4] -= pf * rr[0] * rr[2];
                v6[5] -= pf * rr[1] * rr[2];
            }
        }
        // angle (intra_harm_angle)
        {
            const double dot = r1[0] * r2[0] + r1[1] * r2[1] + r1[2] * r2[2];
            const double r1m1 = 1.0 / roh1, r2m1 = 1.0 / roh2;
            const double r1m2 = r1m1 * r1m1, r2m2 = r2m1 * r2m1, r12 = r1m1 * r2m1;
            const double c = dot * r12;
            const double dxdc = 1.0 / sqrt(1.0 - c * c);
            const double c2r1 = dot * r1m2 * r12, c2r2 = dot * r2m2 * r12;
            const double td = acos(c) - ff.theta_eq;
            en += ff.half_ktheta * td * td;
            const double pf = ff.ktheta * td * dxdc;
            double ga[3], gb[3];  // angle forces on H1, H2
            for (int d = 0; d < 3; ++d) {
                fo[d] += pf * (r12 * (r1[d] + r2[d]) - c2r1 * r1[d] - c2r2 * r2[d]);
                ga[d] = -pf * (r12 * r2[d] - c2r1 * r1[d]);
                gb[d] = -pf * (r12 * r1[d] - c2r2 * r2[d]);
                fh1[d] += ga[d];
                fh2[d] += gb[d];
            }
            if (VIR) {  // sum_i r_i (x) F_i = -(r1 (x) F_H1 + r2 (x) F_H2), r_k = r_O - r_Hk
                v6[0] -= r1[0] * ga[0] + r2[0] * gb[0];
                v6[1] -= r1[1] * ga[1] + r2[1] * gb[1];
                v6[2] -= r1[2] * ga[2] + r2[2] * gb[2];
                v6[3] -= 0.5 * (r1[0] * ga[1] + r2[0] * gb[1] + r1[1] * ga[0] + r2[1] * gb[0]);
                v6[4] -= 0.5 * (r1[0] * ga[2] + r2[0] * gb[2] + r1[2] * ga[0] + r2[2] * gb[0]);
                v6[5] -= 0.5 * (r1[1] * ga[2] + r2[1] * gb[2] + r1[2] * ga[1] + r2[2] * gb[1]);
            }
        }
        double *O = flj + (long)e * 3 * NAP;
        for (int d = 0; d < 3; ++d) {
            O[d * NAP + a] = fo[d];
            O[d * NAP + a + 1] = fh1[d];
            O[d * NAP + a + 2] = fh2[d];
        }
    }
    en = block_sum(en, sh);
    if (threadIdx.x == 0) epart[(long)e * ep_stride + ep_off + mb] = en;
    if (VIR) {
        for (int c = 0; c < 6; ++c) {
            const double v = block_sum(v6[c], sh);
            if (threadIdx.x == 0) vpart[((long)e * 6 + c) * ep_stride + ep_off + mb] = v;
        }
    }
}

// Splined pair potentials of the fast-CMD path (forces.cpp:634-691,
// get_spline :452-466): every atom pair (O-O, O-H, H-H by the atoms'
// slots; intramolecular pairs included) within its table's range, minimum
// image.  Tables are uniform cubic B-splines: coefficients c_{-1..n} of type
// t at coef[off[t] ..], knots a + j h.  Lane <-> atom i of a 64-atom block;
// each lane sums its pairs with every other atom (both orders, no reaction
// writes), the block's 4 waves taking every 4th partner block; energy and
// virial halve the double count.  Writes the full spline force of each atom
// to fspl and the block's energy to epart slot ep_off + I.
// grid: E * nblk, 256 threads.
struct SplTab {
    double a, h, inv_h;
    double cutsq[3];
    int n[3], off[3];
};
__device__ __forceinline__ double spline_eval(const double *__restrict__ c0, int n, double a,
                                              double inv_h, double x, double &dv)
{
    const double t = (x - a) * inv_h;
    int k = (int)floor(t);
    k = max(0, min(k, n - 2));
    const double u = t - k, v = 1.0 - u;
    const double *c = c0 + k;  // c[0] = c_{k-1}
    const double u2 = u * u, u3 = u2 * u;
    const double w0 = v * v * v, w1 = 3.0 * u3 - 6.0 * u2 + 4.0, w2 = -3.0 * u3 + 3.0 * u2 + 3.0 * u + 1.0;
    dv = (c[0] * (-0.5 * v * v) + c[1] * (1.5 * u2 - 2.0 * u) + c[2] * (-1.5 * u2 + u + 0.5) +
          c[3] * (0.5 * u2)) * inv_h;
    return (c[0] * w0 + c[1] * w1 + c[2] * w2 + c[3] * u3) * (1.0 / 6.0);
}
template <bool VIR>
__global__ __launch_bounds__(256) void k_spline(const double *__restrict__ pos, long pstride,
                                                int n_atoms, int NAP, int nblk, Box box, SplTab st,
                                                const double *__restrict__ coef,
                                                double *__restrict__ fspl,
                                                double *__restrict__ epart, int ep_stride,
                                                int ep_off, double *__restrict__ vpart)
{
    const int e = blockIdx.x / nblk, I = blockIdx.x - e * nblk;
    const int l = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const double *X = pos + (long)e * pstride;
    __shared__ double sx[4][64], sy[4][64], sz[4][64];
    __shared__ double red[4][3][64];
    __shared__ double ewv[4], wv[4][6];
    const int i = I * 64 + l;
    const bool vi = i < n_atoms;
    const double xi = vi ? X[i] : 0.0, yi = vi ? X[NAP + i] : 0.0, zi = vi ? X[2 * NAP + i] : 0.0;
    const int oi = (i % 3 == 0);
    double fx = 0.0, fy = 0.0, fz = 0.0, en = 0.0;
    double v6[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int rounds = (nblk + 3) / 4;
    for (int rd = 0; rd < rounds; ++rd) {
        const int J = rd * 4 + w;
        const int jj = J * 64 + l;
        const bool vj = J < nblk && jj < n_atoms;
        sx[w][l] = vj ? X[jj] : 0.0;
        sy[w][l] = vj ? X[NAP + jj] : 0.0;
        sz[w][l] = vj ? X[2 * NAP + jj] : 0.0;
        __syncthreads();
        if (J < nblk) {
            const int cmax = min(64, n_atoms - J * 64);
            for (int c = 0; c < cmax; ++c) {
                const int j = J * 64 + c;
                const double dx = mimg_exact(xi - sx[w][c], box.L[0]);
                const double dy = mimg_exact(yi - sy[w][c], box.L[1]);
                const double dz = mimg_exact(zi - sz[w][c], box.L[2]);
                const double rsq = dx * dx + dy * dy + dz * dz;
                const int oj = (j % 3 == 0);
                const int t = oi ? (oj ? 0 : 1) : (oj ? 1 : 2);
                if (vi && j != i && rsq < st.cutsq[t]) {
                    const double r = sqrt(rsq);
                    double dv;
                    const double v = spline_eval(coef + st.off[t], st.n[t], st.a, st.inv_h, r, dv);
                    const double fp = dv * (1.0 / r);
                    en += 0.5 * v;
                    fx -= fp * dx;
                    fy -= fp * dy;
                    fz -= fp * dz;
                    if (VIR) {
                        const double hp = -0.5 * fp;
                        v6[0] += hp * dx * dx;
                        v6[1] += hp * dy * dy;
                        v6[2] += hp * dz * dz;
                        v6[3] += hp * dx * dy;
                        v6[4] += hp * dx * dz;
                        v6[5] += hp * dy * dz;
                    }
                }
            }
        }
        __syncthreads();
    }
    red[w][0][l] = fx;
    red[w][1][l] = fy;
    red[w][2][l] = fz;
    en = wave_sum(en);
    if (l == 0) ewv[w] = en;
    if (VIR) {
        for (int c = 0; c < 6; ++c) {
            const double v = wave_sum(v6[c]);
            if (l == 0) wv[w][c] = v;
        }
    }
    __syncthreads();
    if (w < 3) {
        const double v = red[0][w][l] + red[1][w][l] + red[2][w][l] + red[3][w][l];
        fspl[((long)e * 3 + w) * NAP + i] = vi ? v : 0.0;
    } else if (l == 0) {
        epart[(long)e * ep_stride + ep_off + I] = ewv[0] + ewv[1] + ewv[2] + ewv[3];
    }
    if (VIR && w == 3 && l < 6)
        vpart[((long)e * 6 + l) * ep_stride + ep_off + I] = wv[0][l] + wv[1][l] + wv[2][l] + wv[3][l];
}

// Virial of E evaluations: the 6 components' partials summed per evaluation.
// grid: E, 64 threads.   out[e*6 + c], c = xx, yy, zz, xy, xz, yz
__global__ __launch_bounds__(64) void k_vsum(const double *__restrict__ vpart, int ep_stride,
                                             int ep_n, double *__restrict__ out)
{
    const int e = blockIdx.x;
    for (int c = 0; c < 6; ++c) {
        const double v = wave_sum_array(vpart + ((long)e * 6 + c) * ep_stride, ep_n);
        if (threadIdx.x == 0) out[(long)e * 6 + c] = v;
    }
}

// Energies of E evaluations (ordered sum of the partials minus the self term).
// grid: 1 block of 64 threads per 64 evaluations.
__global__ __launch_bounds__(64) void k_esum(int E, const double *__restrict__ epart, int ep_stride,
                                             int ep_n, double eself, double *__restrict__ eout)
{
    const int e = blockIdx.x;
    const double v = eval_energy(epart, e, ep_stride, ep_n, eself);
    if (threadIdx.x == 0) eout[e] = v;
}

// ---------------------------------------------------------------------------
// Suzuki-Chin combination (forces.cpp:698-789)
// ---------------------------------------------------------------------------
struct SCArgs {
    int P, n_atoms, NAP;
    double omega_psq, epsilonsc, halfdt;
};

// After the physical-force phase.  grid: P, 256 threads.
//  even b: F_sc = 2 f / 3; tot_b = 2V; vop_b = V
//  odd  b: tot_b = 4V, fsq_b = sum f^2/m / (3 w^2); delta and the displaced
//          positions q +- eps f/m for the finite-difference phase.
// Suzuki-Chin combination, first half (forces.cpp:738-775), one 1024-thread
// block per bead: bead energy; even beads F_sc = 2f/3; odd beads |f|^2_M,
// the finite-difference step eps / sqrt(sum (f/m)^2 / (n_atoms P)) and the
// displaced positions q +- step f/m of the second force phase.
constexpr int SC_THREADS = 1024;
// f2 (may be null): the second half of split k-space forces, added into fph
// first (the sum stays in fph for the later kernels).
__global__ __launch_bounds__(SC_THREADS) void k_sc1(SCArgs a, const double *__restrict__ q,
                                                    double *__restrict__ fph,
                                                    const double *__restrict__ f2,
                                                    const double *__restrict__ epart, int ep_stride,
                                                    int ep_n, double eself,
                                                    double *__restrict__ Vout,
                                                    const double *__restrict__ m_atom,
                                                    double *__restrict__ Fsc,
                                                    double *__restrict__ qpm,
                                                    double *__restrict__ scal /*[P][4]*/)
{
    __shared__ double sh[16];
    __shared__ double s_scale;
    __shared__ double s_V;
    const int b = blockIdx.x;
    const long o = (long)b * 3 * a.NAP;
    if (f2) {
        for (int t = threadIdx.x; t < 3 * a.NAP; t += SC_THREADS) fph[o + t] += f2[o + t];
        __syncthreads();
    }
    if (threadIdx.x < 64) {
        const double v = eval_energy(epart, b, ep_stride, ep_n, eself);
        if (threadIdx.x == 0) {
            s_V = v;
            Vout[b] = v;
        }
    }
    __syncthreads();
    const double V = s_V;
    if ((b & 1) == 0) {
        for (int t = threadIdx.x; t < 3 * a.NAP; t += SC_THREADS) Fsc[o + t] = 2.0 * fph[o + t] / 3.0;
        if (threadIdx.x == 0) {
            scal[b * 4 + 0] = 2.0 * V;
            scal[b * 4 + 1] = V;
            scal[b * 4 + 2] = 0.0;
            scal[b * 4 + 3] = 0.0;
        }
        return;
    }
    double fsq = 0.0, del = 0.0;
    for (int d = 0; d < 3; ++d) {
        for (int at = threadIdx.x; at < a.n_atoms; at += SC_THREADS) {
            const double f = fph[o + d * a.NAP + at];
            const double m = m_atom[at];
            fsq += f * f / m;
            const double fm = f / m;
            del += fm * fm;
        }
    }
    fsq = block_sum(fsq, sh);
    del = block_sum(del, sh);
    if (threadIdx.x == 0) {
        fsq /= 3.0 * a.omega_psq;
        del /= (double)(a.n_atoms * a.P);
        del = 1.0 / sqrt(del);
        const double sc_scale = a.epsilonsc * del;
        scal[b * 4 + 0] = 4.0 * V;
        scal[b * 4 + 1] = 0.0;
        scal[b * 4 + 2] = fsq;
        scal[b * 4 + 3] = 1.0 / (9.0 * a.omega_psq * sc_scale);
        s_scale = sc_scale;
    }
    __syncthreads();
    const double sc_scale = s_scale;
    const int k = b >> 1;
    double *Pp = qpm + (long)(2 * k) * 3 * a.NAP;
    double *Pm = qpm + (long)(2 * k + 1) * 3 * a.NAP;
    for (int d = 0; d < 3; ++d) {
        for (int at = threadIdx.x; at < a.NAP; at += SC_THREADS) {
            const long t = (long)d * a.NAP + at;
            const double dq = sc_scale * fph[o + t] / m_atom[at];
            Pp[t] = q[o + t] + dq;
            Pm[t] = q[o + t] - dq;
        }
    }
}
// Second half (forces.cpp:776-789): odd beads F_sc = 4f/3 - (f+ - f-) / (9 w^2
// step); the half-step kick p += dt/2 F_sc when requested; an extra last
// block sums the SC potential and sc_pot in bead order (beside the others).
// grid: P + 1, SC_THREADS threads.
// f2 (may be null): the second half of the split k-space forces of the
// phase-2 evaluations, added to fpm here.
__global__ __launch_bounds__(SC_THREADS) void k_sc2(SCArgs a, const double *__restrict__ fph,
                                                    const double *__restrict__ fpm,
                                                    const double *__restrict__ f2,
                                                    const double *__restrict__ scal,
                                                    double *__restrict__ Fsc, double *__restrict__ p,
                                                    int kick, double *__restrict__ out,
                                                    unsigned long long *__restrict__ step_ctr)
{
    const int b = blockIdx.x;
    const long o = (long)b * 3 * a.NAP;
    if (b == a.P) {
        if (threadIdx.x == 0) {
            double tot = 0.0, vop = 0.0;
            for (int bb = 0; bb < a.P; ++bb) {
                if ((bb & 1) == 0) {
                    tot += scal[bb * 4 + 0];
                    vop += scal[bb * 4 + 1];
                } else {
                    tot += scal[bb * 4 + 0];
                    tot += scal[bb * 4 + 2];
                }
            }
            tot /= 3.0;
            out[0] = tot;
            out[1] = 2.0 * vop / a.P;
            if (step_ctr) step_ctr[0] += 1ull;
        }
        return;
    }
    if (b & 1) {
        const int k = b >> 1;
        const double fscale = scal[b * 4 + 3];
        const long op = (long)(2 * k) * 3 * a.NAP, om = (long)(2 * k + 1) * 3 * a.NAP;
        for (int t = threadIdx.x; t < 3 * a.NAP; t += SC_THREADS) {
            double F = 4.0 * fph[o + t] / 3.0;
            const double fp = f2 ? fpm[op + t] + f2[op + t] : fpm[op + t];
            const double fm = f2 ? fpm[om + t] + f2[om + t] : fpm[om + t];
            F -= fscale * (fp - fm);
            Fsc[o + t] = F;
            if (kick) p[o + t] += a.halfdt * F;
        }
    } else if (kick) {
        for (int t = threadIdx.x; t < 3 * a.NAP; t += SC_THREADS) p[o + t] += a.halfdt * Fsc[o + t];
    }
}

// ---------------------------------------------------------------------------
// Ring-polymer propagation (update.cpp:281-352), PILE-G (thermostat.cpp:64-133)
// ---------------------------------------------------------------------------
struct PropArgs {
    int P, n_atoms, NAP, ntile;
    double halfdt, dt, beta_n;
    int nvt, host_dev, n_dof;  // host_dev: deviates from the host buffer
    uint64_t seed;
};

// Per-mode constants: mode[k] = {scale_k, wsq_k, c1_k, c2_k}
// (update.cpp:302-304, thermostat.cpp:119-121)

// sqrt-Cayley half step of one normal mode (update.cpp:307-317)
__device__ __forceinline__ void cayley_half(double &uq, double &up, double m, double halfdt,
                                            double4 md)
{
    double temp = halfdt * up / m + uq;
    temp *= md.x;
    up = up - halfdt * md.y * uq * m;
    up *= md.x;
    uq = temp;
}

// Kick, forward normal-mode transform and first half step for mode k of 64
// planar dofs; modes k >= 1 also get their Langevin step and second half
// step here (they do not wait for the centroid).  One wavefront per
// (dof tile, mode); the transform rows are wave-uniform (scalar loads).
// grid: ntile * P (tile-major, XCD-remapped), 64 threads.
// part[tile] = centroid KE after the first half step (tile partial);
// part[ntile + tile] = chi-square partial (device random numbers);
// hpart[k * ntile + tile] = KE before - KE after of mode k's Langevin step.
__global__ __launch_bounds__(64) void k_prop_fwd(PropArgs A, const double *__restrict__ Cm,
                                                 const double4 *__restrict__ mode,
                                                 const double *__restrict__ q,
                                                 const double *__restrict__ p,
                                                 const double *__restrict__ F,
                                                 const double *__restrict__ m_atom,
                                                 const double *__restrict__ sqm_atom,
                                                 const double *__restrict__ dev,
                                                 double *__restrict__ nq, double *__restrict__ np,
                                                 double *__restrict__ part,
                                                 double *__restrict__ hpart,
                                                 const unsigned long long *__restrict__ step_ctr)
{
    const int P = A.P, lb = xcd_remap(blockIdx.x, gridDim.x);
    const int tile = lb / P, k = lb - tile * P, lane = threadIdx.x;
    const long N3 = 3L * A.NAP;
    const int f = tile * 64 + lane;
    const int at = f % A.NAP, d = f / A.NAP;
    const bool va = at < A.n_atoms;
    const double m = va ? m_atom[at] : 1.0;
    const double *cr = Cm + (long)k * P;
    double uq = 0.0, up = 0.0;
#pragma unroll 8
    for (int b = 0; b < P; ++b) {
        const double c = cr[b];
        const long idx = b * N3 + f;
        const double pk = p[idx] + A.halfdt * F[idx];  // update.cpp:281-287
        uq = fma(c, q[idx], uq);
        up = fma(c, pk, up);
    }
    const double4 md = mode[k];
    cayley_half(uq, up, m, A.halfdt, md);
    if (k == 0) {
        nq[f] = uq;
        np[f] = up;
        double ke0 = va ? 0.5 * up * up / m : 0.0;
        double chi = 0.0;
        if (A.nvt && !A.host_dev && va) {
            const int ridx = 3 * at + d;
            if (ridx >= 1) {
                const double z = philox_normal(A.seed, *step_ctr, (uint32_t)(P + 1), (uint32_t)ridx);
                chi = z * z;
            }
        }
        ke0 = wave_sum(ke0);
        chi = wave_sum(chi);
        if (lane == 0) {
            part[tile] = ke0;
            part[A.ntile + tile] = chi;
        }
        return;
    }
    if (A.nvt) {
        double h = 0.0;
        if (va) {
            h += 0.5 * up * up / m;
            const int ridx = 3 * at + d;
            const double xi = A.host_dev ? dev[3 + (long)(k - 1) * A.n_dof + ridx]
                                         : philox_normal(A.seed, *step_ctr, (uint32_t)k, (uint32_t)ridx);
            up = md.z * up + sqm_atom[at] * md.w * xi;
            h -= 0.5 * up * up / m;
        }
        h = wave_sum(h);
        if (lane == 0) hpart[(long)k * A.ntile + tile] = h;
    }
    cayley_half(uq, up, m, A.halfdt, md);
    nq[k * N3 + f] = uq;
    np[k * N3 + f] = up;
}

// Centroid rescale (Bussi, tau -> 0), its second half step, inverse
// transform to bead b and displacement tracking.  One wavefront per
// (dof tile, bead).   grid: ntile * P (tile-major, XCD-remapped), 64 threads.
// hpart[tile] (mode-0 slot) = - centroid KE after the rescale (bead-0 waves).
__global__ __launch_bounds__(64) void k_prop_inv(PropArgs A, const double *__restrict__ Cm,
                                                 const double4 *__restrict__ mode,
                                                 const double *__restrict__ nq,
                                                 const double *__restrict__ np,
                                                 const double *__restrict__ m_atom,
                                                 const double *__restrict__ part,
                                                 const double *__restrict__ dev,
                                                 double *__restrict__ q, double *__restrict__ p,
                                                 double *__restrict__ disp, int NMP,
                                                 double *__restrict__ hpart,
                                                 const unsigned long long *__restrict__ step_ctr)
{
    const int P = A.P, lb = xcd_remap(blockIdx.x, gridDim.x);
    const int tile = lb / P, b = lb - tile * P, lane = threadIdx.x;
    const long N3 = 3L * A.NAP;
    const int f = tile * 64 + lane;
    const int at = f % A.NAP, d = f / A.NAP;
    const bool va = at < A.n_atoms;
    const double m = va ? m_atom[at] : 1.0;
    double u0q = nq[f], u0p = np[f];
    if (A.nvt) {
        const double ke = wave_sum_array(part, A.ntile);
        double R, g2;
        if (A.host_dev) {
            R = dev[0];
            g2 = (A.n_dof % 2 == 0) ? 2.0 * dev[2] + dev[1] * dev[1] : 2.0 * dev[2];
        } else {
            g2 = wave_sum_array(part + A.ntile, A.ntile);
            R = philox_normal(A.seed, *step_ctr, (uint32_t)P, 0u);
        }
        const double alpha = sqrt((R * R + g2) / (2.0 * A.beta_n * ke));
        u0p *= alpha;
        if (b == 0) {
            double h = va ? -0.5 * u0p * u0p / m : 0.0;
            h = wave_sum(h);
            if (lane == 0) hpart[tile] = h;
        }
    }
    cayley_half(u0q, u0p, m, A.halfdt, mode[0]);
    double xq = Cm[b] * u0q, xp = Cm[b] * u0p;
#pragma unroll 8
    for (int k = 1; k < P; ++k) {
        const double c = Cm[(long)k * P + b];
        xq = fma(c, nq[k * N3 + f], xq);
        xp = fma(c, np[k * N3 + f], xp);
    }
    const long idx = b * N3 + f;
    const double qold = q[idx];
    q[idx] = xq;
    p[idx] = xp;
    if (va && (at % 3) == 0) disp[((long)b * 3 + d) * NMP + at / 3] += xq - qold;
}

// The thermostat's heat for one step: pile_g's return value (KE before -
// KE after, thermostat.cpp:86-129); heat_acc -= it (update.cpp:325).
__global__ __launch_bounds__(1024) void k_heat(int ntile, int nh, const double *__restrict__ part,
                                               const double *__restrict__ hpart,
                                               double *__restrict__ heat_acc,
                                               double *__restrict__ heat_step)
{
    __shared__ double sh[16];
    double ke = 0.0, h = 0.0;
    for (int i = threadIdx.x; i < ntile; i += blockDim.x) ke += part[i];
    for (int i = threadIdx.x; i < nh; i += blockDim.x) h += hpart[i];
    ke = block_sum(ke, sh);
    h = block_sum(h, sh);
    if (threadIdx.x == 0) {
        const double ret = ke + h;
        if (heat_step) heat_step[0] = ret;
        if (heat_acc) heat_acc[0] -= ret;
    }
}

// Generic orthonormal transform along the bead axis of a[P][len] (fourier.cpp:12-69).
// grid: ceil(len/64), 256 threads.  fwd: u_k = sum_b C[k][b] x_b, else the transpose.
__global__ __launch_bounds__(256) void k_nm_generic(int P, int len, const double *__restrict__ Cm,
                                                    double *__restrict__ a, int fwd)
{
    extern __shared__ double sm[];
    double *Cs = sm;
    double *X = Cs + P * P;
    const int lane = threadIdx.x & 63, kg = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + lane;
    for (int t = threadIdx.x; t < P * P; t += 256) Cs[t] = Cm[t];
    for (int b = kg; b < P; b += 4) X[b * 64 + lane] = c < len ? a[(long)b * len + c] : 0.0;
    __syncthreads();
    for (int k = kg; k < P; k += 4) {
        double u = 0.0;
        for (int b = 0; b < P; ++b) {
            const double cf = fwd ? Cs[k * P + b] : Cs[b * P + k];
            u = fma(cf, X[b * 64 + lane], u);
        }
        if (c < len) a[(long)k * len + c] = u;
    }
}

// PILE-G on normal-mode momenta in planar layout (drop-in pile_g).
// Pass 1: centroid KE partials.  grid: ntile, 256 threads.
__global__ __launch_bounds__(256) void k_pile_a(int n_atoms, int NAP, int ntile,
                                                const double *__restrict__ pnm,
                                                const double *__restrict__ m_atom,
                                                double *__restrict__ part)
{
    __shared__ double sh[16];
    const int f = blockIdx.x * 256 + threadIdx.x;
    double ke = 0.0;
    if (f < 3 * NAP) {
        const int at = f % NAP;
        if (at < n_atoms) ke = 0.5 * pnm[f] * pnm[f] / m_atom[at];
    }
    ke = block_sum(ke, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = ke;
}

// Pass 2: rescale the centroid, Langevin on modes 1..P-1 with host deviates.
// grid: ntile, 256 threads (same tiling as pass 1).
__global__ __launch_bounds__(256) void k_pile_b(int P, int n_atoms, int NAP, int n_dof, int ntile,
                                                double beta_n, const double4 *__restrict__ mode,
                                                const double *__restrict__ m_atom,
                                                const double *__restrict__ sqm_atom,
                                                const double *__restrict__ part,
                                                const double *__restrict__ dev,
                                                double *__restrict__ pnm,
                                                double *__restrict__ hpart)
{
    __shared__ double sh[16];
    __shared__ double s_alpha;
    if (threadIdx.x == 0) {
        double ke = 0.0;
        for (int t = 0; t < ntile; ++t) ke += part[t];
        const double R = dev[0];
        const double g2 = (n_dof % 2 == 0) ? 2.0 * dev[2] + dev[1] * dev[1] : 2.0 * dev[2];
        s_alpha = sqrt((R * R + g2) / (2.0 * beta_n * ke));
    }
    __syncthreads();
    const int f = blockIdx.x * 256 + threadIdx.x;
    double heat = 0.0;
    if (f < 3 * NAP) {
        const int at = f % NAP, d = f / NAP;
        if (at < n_atoms) {
            const double m = m_atom[at], sq = sqm_atom[at];
            const int ridx = 3 * at + d;
            double up = pnm[f] * s_alpha;
            pnm[f] = up;
            heat -= 0.5 * up * up / m;
            for (int k = 1; k < P; ++k) {
                const double4 md = mode[k];
                const long idx = (long)k * 3 * NAP + f;
                double v = pnm[idx];
                heat += 0.5 * v * v / m;
                v = md.z * v + sq * md.w * dev[3 + (long)(k - 1) * n_dof + ridx];
                heat -= 0.5 * v * v / m;
                pnm[idx] = v;
            }
        }
    }
    heat = block_sum(heat, sh);
    if (threadIdx.x == 0) hpart[blockIdx.x] = heat;
}

// ---------------------------------------------------------------------------
// Verlet lists (neighbour.cpp:12-62) and the rebuild test (update.cpp:356-385)
// ---------------------------------------------------------------------------

// Per bead: two largest O displacement norms; flag = d1 + d2 > r_skin - r_cut.
// grid: P, 256 threads.
__global__ __launch_bounds__(256) void k_check(int n_mols, int NMP, const double *__restrict__ disp,
                                               double skin, int force_all, int *__restrict__ flag,
                                               double *__restrict__ nreb)
{
    __shared__ double s1[256], s2[256];
    const int b = blockIdx.x;
    double m1 = 0.0, m2 = 0.0;
    for (int i = threadIdx.x; i < n_mols; i += 256) {
        const double *D = disp + (long)b * 3 * NMP;
        const double dx = D[i], dy = D[NMP + i], dz = D[2 * NMP + i];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > m1) {
            m2 = m1;
            m1 = d2;
        } else if (d2 > m2) {
            m2 = d2;
        }
    }
    s1[threadIdx.x] = m1;
    s2[threadIdx.x] = m2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            const double a1 = s1[threadIdx.x], a2 = s2[threadIdx.x];
            const double b1 = s1[threadIdx.x + o], b2 = s2[threadIdx.x + o];
            // top two of {a1 >= a2, b1 >= b2}
            double t1, t2;
            if (a1 >= b1) {
                t1 = a1;
                t2 = a2 >= b1 ? a2 : b1;
            } else {
                t1 = b1;
                t2 = b2 >= a1 ? b2 : a1;
            }
            s1[threadIdx.x] = t1;
            s2[threadIdx.x] = t2;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int f = force_all || (sqrt(s1[0]) + sqrt(s2[0]) > skin);
        flag[b] = f;
        if (f && nreb) nreb[b] += 1.0;
    }
}

// Count O-O pairs within r_skin per row (half: j > i; full: j != i).
// The distance test is bit-for-bit neighbour.cpp:42-53.
// grid: P * ceil(n_mols/4), 256 threads (one row per wavefront).
__global__ __launch_bounds__(256) void k_nl_count(int n_mols, int NAP, const double *__restrict__ q,
                                                  Box box, double r_skin_sq,
                                                  const int *__restrict__ flag,
                                                  int *__restrict__ cnt /*[P][2][n_mols]*/)
{
    const int nrb = (n_mols + 3) / 4;
    const int b = blockIdx.x / nrb;
    if (!flag[b]) return;
    const int i = (blockIdx.x - b * nrb) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n_mols) return;
    const double *Q = q + (long)b * 3 * NAP;
    const double xi = Q[3 * i], yi = Q[NAP + 3 * i], zi = Q[2 * NAP + 3 * i];
    int nh = 0, nf = 0;
    for (int j0 = 0; j0 < n_mols; j0 += 64) {
        const int j = j0 + lane;
        bool in = false;
        if (j < n_mols && j != i) {
            const double rx = mimg_exact(xi - Q[3 * j], box.L[0]);
            const double ry = mimg_exact(yi - Q[NAP + 3 * j], box.L[1]);
            const double rz = mimg_exact(zi - Q[2 * NAP + 3 * j], box.L[2]);
            const double rsq = rx * rx + ry * ry + rz * rz;
            in = rsq < r_skin_sq;
        }
        const unsigned long long bf = __ballot(in);
        const unsigned long long bh = __ballot(in && j > i);
        nf += __popcll(bf);
        nh += __popcll(bh);
    }
    if (lane == 0) {
        cnt[((long)b * 2 + 0) * n_mols + i] = nh;
        cnt[((long)b * 2 + 1) * n_mols + i] = nf;
    }
}

// Exclusive scans -> point (reference layout) and fpoint; zero the bead's
// displacements and the unused tail of its half list (neighbour.cpp:23-32).
// grid: P, 1024 threads.
__global__ __launch_bounds__(1024) void k_nl_scan(int n_mols, int list_size, long fcap,
                                                  const int *__restrict__ flag,
                                                  const int *__restrict__ cnt,
                                                  int *__restrict__ point, int *__restrict__ fpoint,
                                                  int *__restrict__ list, double *__restrict__ disp,
                                                  int NMP, int *__restrict__ status)
{
    const int b = blockIdx.x;
    if (!flag[b]) return;
    __shared__ int sh[1024];
    __shared__ int carry[2];
    if (threadIdx.x < 2) carry[threadIdx.x] = 0;
    __syncthreads();
    for (int which = 0; which < 2; ++which) {
        const int *c = cnt + ((long)b * 2 + which) * n_mols;
        int *outp = which == 0 ? point + (long)b * n_mols : fpoint + (long)b * (n_mols + 1);
        for (int i0 = 0; i0 < n_mols; i0 += 1024) {
            const int i = i0 + threadIdx.x;
            const int v = i < n_mols ? c[i] : 0;
            sh[threadIdx.x] = v;
            __syncthreads();
            for (int o = 1; o < 1024; o <<= 1) {
                const int add = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
                __syncthreads();
                sh[threadIdx.x] += add;
                __syncthreads();
            }
            const int incl = sh[threadIdx.x];
            const int base = carry[which];
            if (i < n_mols) outp[i] = base + incl - v;
            __syncthreads();
            if (threadIdx.x == 1023) carry[which] = base + incl;
            __syncthreads();
        }
        if (which == 1 && threadIdx.x == 0) outp[n_mols] = carry[1];
    }
    __syncthreads();
    const int nh = carry[0], nf = carry[1];
    if (threadIdx.x == 0) {
        if (nh > list_size || (long)nf > fcap) atomicOr(status, 1);
    }
    int *L = list + (long)b * list_size;
    for (int k = nh + threadIdx.x; k < list_size; k += 1024) L[k] = 0;
    for (int t = threadIdx.x; t < 3 * NMP; t += 1024) disp[(long)b * 3 * NMP + t] = 0.0;
}

// Write the list entries in ascending j with ballot + prefix popcount.
// grid: as k_nl_count.
__global__ __launch_bounds__(256) void k_nl_fill(int n_mols, int NAP, const double *__restrict__ q,
                                                 Box box, double r_skin_sq, int list_size, long fcap,
                                                 const int *__restrict__ flag,
                                                 const int *__restrict__ point,
                                                 const int *__restrict__ fpoint,
                                                 int *__restrict__ list, int *__restrict__ flist)
{
    const int nrb = (n_mols + 3) / 4;
    const int b = blockIdx.x / nrb;
    if (!flag[b]) return;
    const int i = (blockIdx.x - b * nrb) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n_mols) return;
    const double *Q = q + (long)b * 3 * NAP;
    const double xi = Q[3 * i], yi = Q[NAP + 3 * i], zi = Q[2 * NAP + 3 * i];
    int kh = point[(long)b * n_mols + i];
    long kf = fpoint[(long)b * (n_mols + 1) + i];
    int *L = list + (long)b * list_size;
    int *FL = flist + (long)b * fcap;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int j0 = 0; j0 < n_mols; j0 += 64) {
        const int j = j0 + lane;
        bool in = false;
        if (j < n_mols && j != i) {
            const double rx = mimg_exact(xi - Q[3 * j], box.L[0]);
            const double ry = mimg_exact(yi - Q[NAP + 3 * j], box.L[1]);
            const double rz = mimg_exact(zi - Q[2 * NAP + 3 * j], box.L[2]);
            const double rsq = rx * rx + ry * ry + rz * rz;
            in = rsq < r_skin_sq;
        }
        const bool inh = in && j > i;
        const unsigned long long bf = __ballot(in);
        const unsigned long long bh = __ballot(inh);
        if (in) {
            const long pos = kf + __popcll(bf & below);
            if (pos < fcap) FL[pos] = j;
        }
        if (inh) {
            const int pos = kh + __popcll(bh & below);
            if (pos < list_size) L[pos] = j;
        }
        kf += __popcll(bf);
        kh += __popcll(bh);
    }
}

// Rebuild test (update.cpp:356-385) and, when it fires, the bead's Verlet
// lists (neighbour.cpp:12-62) in one launch: one block of 1024 threads per
// bead.  Row counts live in LDS (n_mols <= NL_MAXMOL), rows are handled one
// per wavefront with ballot + prefix-popcount compaction; the distance test
// is bit-for-bit neighbour.cpp:42-53.   grid: P.
// (a1, a2) <- the two largest of {a1 >= a2} and {b1 >= b2}
__device__ __forceinline__ void top2_merge(double &a1, double &a2, double b1, double b2)
{
    if (a1 >= b1) {
        a2 = a2 >= b1 ? a2 : b1;
    } else {
        a2 = b2 >= a1 ? b2 : a1;
        a1 = b1;
    }
}
constexpr int NL_MAXMOL = 4096;
// Above this size the rebuild runs on the four-kernel path (k_check,
// k_nl_count, k_nl_scan, k_nl_fill: a wavefront per molecule row instead of
// one block per bead), whose rebuild steps are far shorter at thousands of
// molecules; below it the single fused launch costs less per step.
constexpr int NL_FUSED_MAXMOL = 1024;
// Optional heat sum of a thermostatted step (k_heat's work), done by wave 0
// of block 0 here so that it costs no launch of its own.
struct HeatArgs {
    const double *part;  // centroid KE partials (ntile) then chi partials
    const double *hpart; // per-(mode, tile) heat partials (nh)
    double *heat_acc, *heat_step;
    int ntile, nh;
};
__global__ __launch_bounds__(1024) void k_nlist_bead(int n_mols, int NAP, int NMP,
                                                     const double *__restrict__ q, Box box,
                                                     double r_skin_sq, double skin, int force_all,
                                                     int list_size, long fcap,
                                                     double *__restrict__ disp,
                                                     int *__restrict__ flag,
                                                     double *__restrict__ nreb,
                                                     int *__restrict__ point,
                                                     int *__restrict__ fpoint,
                                                     int *__restrict__ list,
                                                     int *__restrict__ flist,
                                                     int *__restrict__ status, HeatArgs heat)
{
    __shared__ double s1[16], s2[16];
    __shared__ int ch[NL_MAXMOL], cf[NL_MAXMOL];
    __shared__ int s_flag, s_tot[2];
    const int b = blockIdx.x, tid = threadIdx.x;
    if (heat.part && b == (int)gridDim.x - 1) {  // the extra last block: the heat sum only
        if (tid < 64) {
            const double ke = wave_sum_array(heat.part, heat.ntile);
            const double h = wave_sum_array(heat.hpart, heat.nh);
            if (tid == 0) {
                const double ret = ke + h;
                heat.heat_step[0] = ret;
                heat.heat_acc[0] -= ret;
            }
        }
        return;
    }
    double *D = disp + (long)b * 3 * NMP;
    // --- two largest O displacements
    double m1 = 0.0, m2 = 0.0;
    for (int i = tid; i < n_mols; i += 1024) {
        const double dx = D[i], dy = D[NMP + i], dz = D[2 * NMP + i];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > m1) {
            m2 = m1;
            m1 = d2;
        } else if (d2 > m2) {
            m2 = d2;
        }
    }
    // top two of the block: within each wave by shuffles, then over the 16 waves
    for (int o = 32; o > 0; o >>= 1) {
        const double b1 = __shfl_xor(m1, o, 64), b2 = __shfl_xor(m2, o, 64);
        top2_merge(m1, m2, b1, b2);
    }
    if ((tid & 63) == 0) {
        s1[tid >> 6] = m1;
        s2[tid >> 6] = m2;
    }
    __syncthreads();
    if (tid == 0) {
        double t1 = s1[0], t2 = s2[0];
        for (int ww = 1; ww < 16; ++ww) top2_merge(t1, t2, s1[ww], s2[ww]);
        const int f = force_all || (sqrt(t1) + sqrt(t2) > skin);
        s_flag = f;
        flag[b] = f;
        if (f && nreb) nreb[b] += 1.0;
    }
    __syncthreads();
    if (!s_flag) return;
    // --- counts
    const double *Q = q + (long)b * 3 * NAP;
    const int lane = tid & 63, w = tid >> 6;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int i = w; i < n_mols; i += 16) {
        const double xi = Q[3 * i], yi = Q[NAP + 3 * i], zi = Q[2 * NAP + 3 * i];
        int nh = 0, nf = 0;
        for (int j0 = 0; j0 < n_mols; j0 += 64) {
            const int j = j0 + lane;
            bool in = false;
            if (j < n_mols && j != i) {
                const double rx = mimg_exact(xi - Q[3 * j], box.L[0]);
                const double ry = mimg_exact(yi - Q[NAP + 3 * j], box.L[1]);
                const double rz = mimg_exact(zi - Q[2 * NAP + 3 * j], box.L[2]);
                in = rx * rx + ry * ry + rz * rz < r_skin_sq;
            }
            nf += __popcll(__ballot(in));
            nh += __popcll(__ballot(in && j > i));
        }
        if (lane == 0) {
            ch[i] = nh;
            cf[i] = nf;
        }
    }
    __syncthreads();
    // --- exclusive scans (wave 0: chunks of 64 with a running carry)
    if (w == 0) {
        int carry_h = 0, carry_f = 0;
        for (int i0 = 0; i0 < n_mols; i0 += 64) {
            const int i = i0 + lane;
            const int vh = i < n_mols ? ch[i] : 0, vf = i < n_mols ? cf[i] : 0;
            int ih = vh, iff = vf;
            for (int o = 1; o < 64; o <<= 1) {
                const int th = __shfl_up(ih, o, 64), tf = __shfl_up(iff, o, 64);
                if (lane >= o) {
                    ih += th;
                    iff += tf;
                }
            }
            if (i < n_mols) {
                ch[i] = carry_h + ih - vh;
                cf[i] = carry_f + iff - vf;
            }
            carry_h += __shfl(ih, 63, 64);
            carry_f += __shfl(iff, 63, 64);
        }
        if (lane == 0) {
            s_tot[0] = carry_h;
            s_tot[1] = carry_f;
        }
    }
    __syncthreads();
    const int nh_tot = s_tot[0], nf_tot = s_tot[1];
    if (tid == 0 && (nh_tot > list_size || (long)nf_tot > fcap)) atomicOr(status, 1);
    for (int i = tid; i < n_mols; i += 1024) {
        point[(long)b * n_mols + i] = ch[i];
        fpoint[(long)b * (n_mols + 1) + i] = cf[i];
    }
    if (tid == 0) fpoint[(long)b * (n_mols + 1) + n_mols] = nf_tot;
    int *L = list + (long)b * list_size;
    int *FL = flist + (long)b * fcap;
    for (int k = nh_tot + tid; k < list_size; k += 1024) L[k] = 0;
    for (int t = tid; t < 3 * NMP; t += 1024) D[t] = 0.0;
    // --- fill
    for (int i = w; i < n_mols; i += 16) {
        const double xi = Q[3 * i], yi = Q[NAP + 3 * i], zi = Q[2 * NAP + 3 * i];
        int kh = ch[i];
        long kf = cf[i];
        for (int j0 = 0; j0 < n_mols; j0 += 64) {
            const int j = j0 + lane;
            bool in = false;
            if (j < n_mols && j != i) {
                const double rx = mimg_exact(xi - Q[3 * j], box.L[0]);
                const double ry = mimg_exact(yi - Q[NAP + 3 * j], box.L[1]);
                const double rz = mimg_exact(zi - Q[2 * NAP + 3 * j], box.L[2]);
                in = rx * rx + ry * ry + rz * rz < r_skin_sq;
            }
            const bool inh = in && j > i;
            const unsigned long long bf = __ballot(in), bh = __ballot(inh);
            if (in) {
                const long pos = kf + __popcll(bf & below);
                if (pos < fcap) FL[pos] = j;
            }
            if (inh) {
                const int pos = kh + __popcll(bh & below);
                if (pos < list_size) L[pos] = j;
            }
            kf += __popcll(bf);
            kh += __popcll(bh);
        }
    }
}

// ---------------------------------------------------------------------------
// Classical velocity Verlet (update.cpp:78-137)
// ---------------------------------------------------------------------------
// p += dt/2 F; q += p dt / m; O displacements.  grid: ceil(3 NAP / 256).
__global__ __launch_bounds__(256) void k_vv_a(int n_atoms, int NAP, int NMP, double halfdt,
                                              double dt, const double *__restrict__ m_atom,
                                              const double *__restrict__ F, double *__restrict__ q,
                                              double *__restrict__ p, double *__restrict__ disp)
{
    const int f = blockIdx.x * 256 + threadIdx.x;
    if (f >= 3 * NAP) return;
    const int at = f % NAP, d = f / NAP;
    if (at >= n_atoms) return;
    double pp = p[f] + halfdt * F[f];
    p[f] = pp;
    const double m = m_atom[at];
    q[f] += pp * dt / m;
    if (at % 3 == 0) disp[d * NMP + at / 3] += pp * dt / m;
}

// F = f_phys of the single evaluation, p += dt/2 F, energy out.
__global__ __launch_bounds__(256) void k_vv_b(int n_atoms, int NAP, double halfdt,
                                              const double *__restrict__ fe,
                                              const double *__restrict__ epart, int ep_stride,
                                              int ep_n, double eself,
                                              double *__restrict__ F, double *__restrict__ p,
                                              double *__restrict__ out,
                                              unsigned long long *__restrict__ step_ctr)
{
    const int f = blockIdx.x * 256 + threadIdx.x;
    if (f < 3 * NAP) {
        const double v = fe[f];
        F[f] = v;
        const int at = f % NAP;
        if (at < n_atoms) p[f] += halfdt * v;
    }
    if (blockIdx.x == 0 && threadIdx.x < 64) {
        const double v = eval_energy(epart, 0, ep_stride, ep_n, eself);
        if (threadIdx.x == 0) {
            out[0] = v;
            out[1] = 0.0;
            if (step_ctr) step_ctr[0] += 1ull;
        }
    }
}

// ---------------------------------------------------------------------------
// Observables (path_integrals.cpp:10-81)
// ---------------------------------------------------------------------------
// get_kcv: 0.5 n_dof / beta + (1/P) sum_{even b} sum_j -(q_bj - qc_j) f_bj.
// grid: ceil(3 NAP / 256); out: partials per block.
// Observables of the resident state, per 256-dof block: the centroid-virial
// sum -(q_b - qbar).f_b over even beads (path_integrals.cpp:29-48; centroid
// divided by float(P) as :21), the kinetic energy sum p^2/2m and the ring
// polymer spring energy (1/2) m w_P^2 sum_b (q_{b+1} - q_b)^2, which equals
// get_conserved's normal-mode sum (1/2) m w_k^2 q~_k^2 (path_integrals.cpp:
// 63-76) because the transform is orthonormal and w_k = 2 w_P sin(k pi / P).
// part[blk], part[nb + blk], part[2 nb + blk].   grid: nb, 256 threads.
__global__ __launch_bounds__(256) void k_obs_part(int P, int n_atoms, int NAP, double omega_psq,
                                                  const double *__restrict__ q,
                                                  const double *__restrict__ p,
                                                  const double *__restrict__ f,
                                                  const double *__restrict__ m_atom,
                                                  double *__restrict__ part)
{
    __shared__ double sh[16];
    const int t = blockIdx.x * 256 + threadIdx.x;
    double kc = 0.0, ke = 0.0, sp = 0.0;
    if (t < 3 * NAP && (t % NAP) < n_atoms) {
        const long N3 = 3L * NAP;
        const double m = m_atom[t % NAP];
        double qc = 0.0;
        for (int b = 0; b < P; ++b) qc += q[b * N3 + t];
        qc /= (double)(float)P;
        for (int b = 0; b < P; ++b) {
            const double qb = q[b * N3 + t], pb = p[b * N3 + t];
            if ((b & 1) == 0) kc -= (qb - qc) * f[b * N3 + t];
            ke += 0.5 * pb * pb / m;
            if (P > 1) {
                const double dq = q[((b + 1) % P) * N3 + t] - qb;
                sp += 0.5 * m * omega_psq * dq * dq;
            }
        }
    }
    kc = block_sum(kc, sh);
    ke = block_sum(ke, sh);
    sp = block_sum(sp, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = kc;
        part[gridDim.x + blockIdx.x] = ke;
        part[2 * gridDim.x + blockIdx.x] = sp;
    }
}

// Reduce k_obs_part's partials into {K_cv, conserved, energy, sc_pot} (P > 1)
// or {kinetic energy, V + KE, V, 0} (P = 1).  With a trace, only on steps
// whose counter is a multiple of `every`, into trace[n][4], n counted in
// *ntrace (host-cleared before a run), at most cap samples.
// grid: 1, 64 threads.
__global__ __launch_bounds__(64) void k_obs_fin(int P, int nb, double half_ndof_over_beta,
                                                const double *__restrict__ part,
                                                const double *__restrict__ out,
                                                const unsigned long long *__restrict__ step_ctr,
                                                int every, int cap, int *__restrict__ ntrace,
                                                double *__restrict__ dst)
{
    if (every > 0 && (step_ctr[0] % (unsigned long long)every) != 0ull) return;
    const double kc = wave_sum_array(part, nb);
    const double ke = wave_sum_array(part + nb, nb);
    const double sp = wave_sum_array(part + 2 * nb, nb);
    if (threadIdx.x == 0) {
        double *o = dst;
        if (every > 0) {
            const int n = ntrace[0];
            if (n >= cap) return;
            o = dst + 4L * n;
            ntrace[0] = n + 1;
        }
        const double e = out[0];
        if (P > 1) {
            o[0] = half_ndof_over_beta + kc / P;
            o[1] = e + ke + sp;
            o[2] = e;
            o[3] = out[1];
        } else {
            o[0] = ke;
            o[1] = e + ke;
            o[2] = e;
            o[3] = 0.0;
        }
    }
}

__global__ __launch_bounds__(256) void k_kcv(int P, int n_atoms, int NAP,
                                             const double *__restrict__ q,
                                             const double *__restrict__ f,
                                             double *__restrict__ part)
{
    __shared__ double sh[16];
    const int t = blockIdx.x * 256 + threadIdx.x;
    double acc = 0.0;
    if (t < 3 * NAP && (t % NAP) < n_atoms) {
        const long N3 = 3L * NAP;
        double qc = 0.0;
        for (int b = 0; b < P; ++b) qc += q[b * N3 + t];
        qc /= (double)(float)P;
        for (int b = 0; b < P; b += 2) acc -= (q[b * N3 + t] - qc) * f[b * N3 + t];
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// get_conserved on normal-mode q, p: sum_k sum_j p^2/2m + m w_k^2 q^2 / 2.
__global__ __launch_bounds__(256) void k_cons(int P, int n_atoms, int NAP,
                                              const double *__restrict__ nq,
                                              const double *__restrict__ np,
                                              const double4 *__restrict__ mode,
                                              const double *__restrict__ m_atom,
                                              double *__restrict__ part)
{
    __shared__ double sh[16];
    const int t = blockIdx.x * 256 + threadIdx.x;
    double acc = 0.0;
    if (t < 3 * NAP && (t % NAP) < n_atoms) {
        const long N3 = 3L * NAP;
        const double m = m_atom[t % NAP];
        for (int k = 0; k < P; ++k) {
            const double pv = np[k * N3 + t], qv = nq[k * N3 + t];
            acc += 0.5 * pv * pv / m + 0.5 * mode[k].y * (m * qv * qv);
        }
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

}  // namespace qtp
