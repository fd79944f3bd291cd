// qtp_kspace.hpp -- reciprocal-space Ewald (forces.cpp:180-448): structure factors,
// their reduction, forces; matrix-core and vector forms.
#pragma once
#include "qtp_realspace.hpp"

namespace qtp {
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
// forms its A operand F_p with one product of two table entries.
// A block takes cpb consecutive chunks (a chunk group) and a slice of the
// units: it builds each chunk's factor tables in turn and keeps its units'
// accumulators in registers across the chunks (at most KS_MAXU units per
// wave), so one partial per group and chunk group is written: cpb times
// fewer partials to write here and to read in k_sreduce.
// Output per group and chunk group: Re/Im Z times F0..F3, in that order.
// kcol[c] = (kx | ky << 16, first group, first kz, last kz) per sorted
// column (padding columns: first kz > last kz).  usplit blocks share a chunk
// group, each taking 1/usplit of the units (the host makes the slices fit
// KS_MAXU units per wave).
// grid: E * ncg * usplit (ncg = ceil(nac / cpb)), 64 * KS_WAVES threads,
// ksum_mf_lds_bytes(K, AC, ncol4, nunit).
constexpr int KS_WAVES = 8;
constexpr int KS_MAXU = 6;  // units per wave of one block (accumulator registers)
// LDS rows of AC + 2 entries: the B operands of a quarter-wave (16 rows n at
// one site) and the A factors of a tile's columns then fall in different banks.
// The column and unit tables are staged in LDS behind the factor tables.
// z-factor rows of k_ksum_mf: 2 K entries padded to whole 16-row n tiles
// (at least two)
__host__ __device__ inline int ksum_mf_zrows(int K) { return 2 * K <= 32 ? 32 : ((2 * K + 15) / 16) * 16; }
inline size_t ksum_mf_lds_bytes(int K, int AC, int ncol4, int nunit)
{
    return ((size_t)4 * K + ksum_mf_zrows(K)) * (AC + 2) * 8 + (size_t)ncol4 * 16 +
           (size_t)nunit * 8;
}

__global__ __launch_bounds__(64 * KS_WAVES) void k_ksum_mf(
    const double *__restrict__ pos, long pstride, const double *__restrict__ charge, int n_atoms,
    int NAP, double gamma, double half1mg, Ewald ew, const int4 *__restrict__ kcol, int ncol4,
    const int2 *__restrict__ kunit, int nunit, int G, int AC, int nac, int cpb, int usplit,
    double *__restrict__ spart)
{
    extern __shared__ double smem[];
    const int K = ew.K;
    const int RS = AC + 2;           // LDS row stride
    const int ZN = ksum_mf_zrows(K);
    double2 *T = (double2 *)smem;    // [2][K][RS] x (charge folded), y factors
    double *Zt = smem + 4 * K * RS;  // [ZN][RS]
    int4 *sc = (int4 *)(Zt + ZN * RS);  // [ncol4] column table
    int2 *su = (int2 *)(sc + ncol4);    // [nunit] unit table
    const int ncg = (nac + cpb - 1) / cpb;
    const int per_e = ncg * usplit;
    const int e = blockIdx.x / per_e, r = blockIdx.x - e * per_e;
    const int cg = r / usplit, us = r - cg * usplit;
    const int u0 = us * nunit / usplit, u1 = (us + 1) * nunit / usplit;
    const int ac0 = cg * cpb, ac1 = min(nac, ac0 + cpb);
    const double *X = pos + (long)e * pstride;
    for (int t = threadIdx.x; t < ncol4; t += blockDim.x) sc[t] = kcol[t];
    for (int t = threadIdx.x; t < nunit; t += blockDim.x) su[t] = kunit[t];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int j16 = lane & 15, gq = lane >> 4;
    // A row j16 = (column j16 / 4, product p = lane % 4): F_p = X_a Y_b with
    // a = Re for p = 0, 2, b = Re for p = 0, 3 (Im otherwise)
    const int pp = lane & 3;
    const int xa = (pp == 1 || pp == 3) ? 1 : 0, yb = (pp == 1 || pp == 2) ? 1 : 0;
    const double *Td = (const double *)T;
    mf_v4d acc[KS_MAXU];
#pragma unroll
    for (int k = 0; k < KS_MAXU; ++k) acc[k] = mf_v4d{0.0, 0.0, 0.0, 0.0};
    for (int ac = ac0; ac < ac1; ++ac) {
        if (ac > ac0) __syncthreads();  // the previous chunk's tables are read
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
                for (int n = 2 * K; n < ZN; ++n) Zt[n * RS + j] = 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KS_MAXU; ++k) {
            const int u = u0 + w + k * KS_WAVES;
            if (u >= u1) break;
            const int2 un = su[u];
            const int ck = sc[4 * un.x + (j16 >> 2)].x;
            const int kx = ck & 0xffff, ky = ck >> 16;
            const double *Xr = Td + 2 * (kx * RS + gq) + xa;  // site 4 s + gq
            const double *Yr = Td + 2 * ((K + ky) * RS + gq) + yb;
            const double *Zr = Zt + (16 * un.y + j16) * RS + gq;
            mf_v4d a4 = acc[k];
#pragma unroll 4
            for (int s4 = 0; s4 < AC; s4 += 4)
                a4 = __builtin_amdgcn_mfma_f64_16x16x4f64(Xr[2 * s4] * Yr[2 * s4], Zr[s4], a4, 0, 0, 0);
            acc[k] = a4;
        }
    }
    double *out = spart + ((long)e * ncg + cg) * G * 8;
#pragma unroll
    for (int k = 0; k < KS_MAXU; ++k) {
        const int u = u0 + w + k * KS_WAVES;
        if (u >= u1) break;
        const int2 un = su[u];
        // lane: n = 16 n-tile + j16, product gq, column 4 tile + reg
        const int n = 16 * un.y + j16, kz = n >> 1;
#pragma unroll
        for (int reg = 0; reg < 4; ++reg) {
            const int4 cr = sc[4 * un.x + reg];
            if (kz >= cr.z && kz <= cr.w)
                out[(long)(cr.y + kz - cr.z) * 8 + gq * 2 + (n & 1)] = acc[k][reg];
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
    int chunk, const unsigned char *__restrict__ rfar, int ntile)
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
    // skipped tiles' slots (culling) are not read, as in k_kforce_mf
    const int nsl = (nblk - w + KF_WAVES - 1) / KF_WAVES;
    const bool masked = rfar && half == 0 && nsl <= 64;
    const unsigned char *far = rfar ? rfar + (long)e * ntile : nullptr;
    const int Bs = atom >> 6;  // the lane's site's 64-site block
    unsigned long long live = 0ull, own = 0ull;
    if (masked) {
        const int a0 = ab * KF_SITES;
        const int B0 = a0 >> 6, B1 = (min(a0 + KF_SITES, n_atoms) - 1) >> 6;
        bool l0 = false, l1 = false;
        if (lane < nsl) {
            const int sl = w + KF_WAVES * lane;
            l0 = !far[tile_index(min(B0, sl), max(B0, sl), nblk)];
            l1 = !far[tile_index(min(B1, sl), max(B1, sl), nblk)];
        }
        const unsigned long long m0 = __ballot(l0), m1 = __ballot(l1);
        live = m0 | m1;
        own = Bs == B0 ? m0 : m1;
    }
    if (va && half == 0) {
        if (masked) {
            for (unsigned long long m = live; m; m &= m - 1ull) {
                const int k = __ffsll((long long)m) - 1;
                if (!((own >> k) & 1ull)) continue;
                const int sl = w + KF_WAVES * k;
                const double *rp = rpart + ((long)e * nblk + sl) * 3 * NAP + atom;
                fx += rp[0];
                fy += rp[NAP];
                fz += rp[2 * NAP];
            }
        } else {
            for (int sl = w; sl < nblk; sl += KF_WAVES) {
                if (far && far[tile_index(min(Bs, sl), max(Bs, sl), nblk)]) continue;
                const double *rp = rpart + ((long)e * nblk + sl) * 3 * NAP + atom;
                fx += rp[0];
                fy += rp[NAP];
                fz += rp[2 * NAP];
            }
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
// k-steps of 4 z-factor entries per tile (A operands in registers): a
// kernel built for MAXKS k-steps takes 2 K <= 4 MAXKS (k_max <= 2 MAXKS - 1);
// MF_KS_LO (k_max <= 15) keeps the register use of the common sizes low,
// MF_MAXKS (k_max <= 23) serves the larger k-sets.
constexpr int MF_KS_LO = 8;
constexpr int MF_MAXKS = 12;

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

template <int W, int MAXKS>
__global__ __launch_bounds__(64 * W) void k_kforce_mf(
    const double *__restrict__ pos, long pstride, const double *__restrict__ charge, int n_atoms,
    int NAP, double gamma, double half1mg, Ewald ew, const double *__restrict__ wm, int Mp,
    int nks, const int4 *__restrict__ rt, int nrt, const double *__restrict__ rpart, int nblk,
    const double *__restrict__ flj, const double *__restrict__ fspl, double *__restrict__ fout,
    const unsigned char *__restrict__ rfar, int ntile)
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
    // real-space slots w, w + W, ...: with culling (rfar), the entries of a
    // skipped tile are never written, so a site reads slot sl only if tile
    // (its 64-site block, sl) was evaluated.  The wave's sites fall in at most
    // two blocks, B0 and B1: one mask per block over the wave's slots, the
    // loop over their union, each lane's loads under its own block's mask.
    const int nsl = (nblk - w + W - 1) / W;
    unsigned long long live = 0ull, own = 0ull;
    const bool masked = rfar && nsl <= 64;
    const unsigned char *far = rfar ? rfar + (long)e * ntile : nullptr;
    const int Bs = (a0 + lane) >> 6;
    if (masked) {
        const int B0 = a0 >> 6, B1 = (min(a0 + MF_SITES, n_atoms) - 1) >> 6;
        bool l0 = false, l1 = false;
        if (lane < nsl) {
            const int sl = w + W * lane;
            l0 = !far[tile_index(min(B0, sl), max(B0, sl), nblk)];
            l1 = !far[tile_index(min(B1, sl), max(B1, sl), nblk)];
        }
        const unsigned long long m0 = __ballot(l0), m1 = __ballot(l1);
        live = m0 | m1;
        own = Bs == B0 ? m0 : m1;
    }
    if (lane < MF_SITES && a0 + lane < n_atoms) {
        if (masked) {
            for (unsigned long long m = live; m; m &= m - 1ull) {
                const int k = __ffsll((long long)m) - 1;
                if (!((own >> k) & 1ull)) continue;
                const int sl = w + W * k;
                const double *rp = rpart + ((long)e * nblk + sl) * 3 * NAP + a0 + lane;
                rx += rp[0];
                ry += rp[NAP];
                rz += rp[2 * NAP];
            }
        } else {
            for (int sl = w; sl < nblk; sl += W) {
                if (far && far[tile_index(min(Bs, sl), max(Bs, sl), nblk)]) continue;
                const double *rp = rpart + ((long)e * nblk + sl) * 3 * NAP + a0 + lane;
                rx += rp[0];
                ry += rp[NAP];
                rz += rp[2 * NAP];
            }
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
        if constexpr (MAXKS <= MF_KS_LO) {
            switch (r.z) {  // 1 <= r.z <= MAXKS (host table)
            case 1: mf_tile<1>(Wr, Mp4, Zl, d0, d1, d2); break;
            case 2: mf_tile<2>(Wr, Mp4, Zl, d0, d1, d2); break;
            case 3: mf_tile<3>(Wr, Mp4, Zl, d0, d1, d2); break;
            case 4: mf_tile<4>(Wr, Mp4, Zl, d0, d1, d2); break;
            case 5: mf_tile<5>(Wr, Mp4, Zl, d0, d1, d2); break;
            case 6: mf_tile<6>(Wr, Mp4, Zl, d0, d1, d2); break;
            case 7: mf_tile<7>(Wr, Mp4, Zl, d0, d1, d2); break;
            default: mf_tile<8>(Wr, Mp4, Zl, d0, d1, d2); break;
            }
        } else {
            // larger k-sets: tiles of up to 12 k-steps, as two straight-line
            // halves (registers of one half at a time)
            const int lo = r.z < 6 ? r.z : 6;
            switch (lo) {
            case 1: mf_tile<1>(Wr, Mp4, Zl, d0, d1, d2); break;
            case 2: mf_tile<2>(Wr, Mp4, Zl, d0, d1, d2); break;
            case 3: mf_tile<3>(Wr, Mp4, Zl, d0, d1, d2); break;
            case 4: mf_tile<4>(Wr, Mp4, Zl, d0, d1, d2); break;
            case 5: mf_tile<5>(Wr, Mp4, Zl, d0, d1, d2); break;
            default: mf_tile<6>(Wr, Mp4, Zl, d0, d1, d2); break;
            }
            const double *Wr2 = Wr + 6 * Mp4;
            const double *Zl2 = Zl + 4 * 6 * MF_SITES;
            switch (r.z - lo) {
            case 0: break;
            case 1: mf_tile<1>(Wr2, Mp4, Zl2, d0, d1, d2); break;
            case 2: mf_tile<2>(Wr2, Mp4, Zl2, d0, d1, d2); break;
            case 3: mf_tile<3>(Wr2, Mp4, Zl2, d0, d1, d2); break;
            case 4: mf_tile<4>(Wr2, Mp4, Zl2, d0, d1, d2); break;
            case 5: mf_tile<5>(Wr2, Mp4, Zl2, d0, d1, d2); break;
            default: mf_tile<6>(Wr2, Mp4, Zl2, d0, d1, d2); break;
            }
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

}  // namespace qtp
