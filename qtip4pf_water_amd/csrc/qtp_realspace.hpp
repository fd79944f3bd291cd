// qtp_realspace.hpp -- real-space Ewald (forces.cpp:141-177, 586-608).
#pragma once
#include "qtp_common.hpp"

namespace qtp {
// Real-space Ewald over all site pairs i<j within r_c (forces.cpp:586-608,
// rwald :141-177).  One block of W wavefronts (W = 2 or 4) per 64x64 tile
// (I<=J) of the site pair matrix; wave w takes the 64/W j-columns
// c0 = w 64/W ... c0 + 64/W - 1 against all 64 i.
// Pass 1 (lane <-> i) marks each lane's in-cutoff columns (a bit mask).
// The wave then lays its in-cutoff pairs out in an LDS queue column by
// column (a ballot per column; within a column, rows ascending), and pass 2
// gives every lane the same number of consecutive queue entries: lanes stay
// busy however unevenly the pairs fall.  Consecutive entries mostly share
// the column j, so a lane keeps the j reaction in registers and adds it to
// the LDS column array when j changes; i forces go to the wave's own LDS
// rows with ds_add_f64.  Columns belong to one wave and every wave has its
// own rows, and the LDS adds of a wave follow program and lane order, so
// the result is reproducible.
// Partial site forces go to slot J (block I's sites) and slot I (block J's
// sites): every (site, slot) has exactly one writer.
// grid: E * nt, 64 W threads: tiles t0 ... t0 + nt - 1 of the ntile of the
// list (a launch may take a part of the list; the parts write disjoint
// slots and partials).
// VIR: also accumulate the pair virial r_ij (x) f_ij (6 components) into
// vpart[(e*6 + c)*ep_stride + ep_off + tile].
// CARG: the special functions' coefficients come from the argument cf (SGPRs)
// instead of literals (VGPRs); the host picks it for large grids.
// CULL: the sites come from k_sites (x, y, z, q per site and bounding boxes
// per 64-site block and per 8-site cluster).  A tile whose two blocks' boxes
// are more than r_c apart under every periodic image holds no pair within the
// cutoff: k_tilefar marks it in rfar[e * ntile + tile] and zeroes its energy
// partial, and its block here returns after reading the flag, writing no slot
// entries (k_kforce_mf and k_kforce read none of a tile so marked).  In the
// other tiles the cutoff test runs only over pairs of clusters whose boxes
// lie within r_c (4096 H2O, r_c 9 A: a third of the cluster pairs of the near
// tiles), one lane per site pair.  The same pairs are evaluated either way.
// The host picks it for boxes several r_c wide, where most tiles are far
// (4096 H2O, r_c 9 A: 76% of the tiles, sites in molecule order).
// W: two waves per tile halve the block and double each lane's share of the
// queue; a launch then holds fewer waves (216 H2O x 32 beads: 4224, which
// the chip holds at once, against 8448 in two rounds with four).
constexpr int RS_WAVES = 4;  // the default W
constexpr long RS_CARG_BLOCKS = 16384;  // grids at least this large use CARG
// boxes at least this many r_c wide use CULL.  x 32 beads, steps/s of the
// first culling form (block boxes only, zeroed slots) against the plain one
// at the reference's Ewald point: 216 H2O (2.06 r_c) -8%, 1000 (2.66) +7%,
// 1728 (2.91) +7%, 4096 (3.36) +21%
constexpr double RS_CULL_MIN_BOX = 2.5;

// Smallest distance between [a0, a1] and the periodic images of [b0, b1]
// (period L): the nearest image on the right is f = (b0 - a1) mod L away,
// the one a period to its left L - f - |A| - |B|; negative means overlap.
// Rounding of the mod only ever lowers the result to 0.
__device__ __forceinline__ double periodic_gap(double a0, double a1, double b0, double b1,
                                               double L, double invL)
{
    double f = b0 - a1;
    f = fma(-L, floor(f * invL), f);
    const double back = L - f - (a1 - a0) - (b1 - b0);
    return fmax(0.0, fmin(f, back));
}

// Index of tile (I, J), I <= J, in the real-space tile list (I-major upper
// triangle of the nblk x nblk block matrix).
__device__ __forceinline__ int tile_index(int I, int J, int nblk)
{
    return I * nblk - (I * (I - 1)) / 2 + (J - I);
}

// Doubles per block in the CULL form's box table: the block's bounding box
// (lo xyz, hi xyz) at [0, 6), then the boxes of its eight 8-site clusters,
// component c of cluster k at [6 + 8 c + k].
constexpr int RS_BOX = 6 + 6 * 8;

// Sites of each 64-site block as (x, y, z, q) (M from O, H, H; forces.cpp:
// 558-564), the block's bounding box and those of its 8-site clusters
// (RS_BOX layout), for the CULL form of k_realspace.  A cluster holding no
// site gets an empty box (lo > hi); k_realspace never tests one.
// grid: E * nblk, 64 threads.
__global__ __launch_bounds__(64) void k_sites(const double *__restrict__ pos, long pstride,
                                              const double *__restrict__ charge, int n_atoms,
                                              int NAP, int nblk, double gamma, double half1mg,
                                              double4 *__restrict__ sites,
                                              double *__restrict__ bbox)
{
    const int e = blockIdx.x / nblk, B = blockIdx.x - e * nblk;
    const int l = threadIdx.x;
    const int s = B * 64 + l;
    const double *X = pos + (long)e * pstride;
    const bool ok = s < n_atoms;  // lane 0 always holds a site
    double4 v = make_double4(0.0, 0.0, 0.0, 0.0);
    if (ok)
        v = make_double4(site_coord(X, NAP, s, 0, gamma, half1mg),
                         site_coord(X, NAP, s, 1, gamma, half1mg),
                         site_coord(X, NAP, s, 2, gamma, half1mg), charge[s]);
    sites[(long)e * nblk * 64 + s] = v;
    double lo[3] = {ok ? v.x : 1e300, ok ? v.y : 1e300, ok ? v.z : 1e300};
    double hi[3] = {ok ? v.x : -1e300, ok ? v.y : -1e300, ok ? v.z : -1e300};
    double *bb = bbox + ((long)e * nblk + B) * RS_BOX;
    // lanes 8k ... 8k + 7 form cluster k: the first three butterfly steps
    // stay inside it
    for (int o = 1; o < 8; o <<= 1)
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], o, 64));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], o, 64));
        }
    {
        const int c = l & 7, k = l >> 3;
        if (c < 6) {
            const double vc = c == 0 ? lo[0] : c == 1 ? lo[1] : c == 2 ? lo[2]
                            : c == 3 ? hi[0] : c == 4 ? hi[1] : hi[2];
            bb[6 + 8 * c + k] = vc;
        }
    }
    for (int o = 8; o < 64; o <<= 1)
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], o, 64));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], o, 64));
        }
    if (l < 3) bb[l] = l == 0 ? lo[0] : (l == 1 ? lo[1] : lo[2]);
    else if (l < 6) bb[l] = l == 3 ? hi[0] : (l == 4 ? hi[1] : hi[2]);
}

// Skip flags of the CULL form, one thread per (evaluation, tile): tile (I, J)
// is skipped (rfar = 1) when its two blocks' boxes are more than r_c apart
// under every periodic image; its energy partial (and virial partials, when
// vpart is given) are zeroed here, so its k_realspace block only reads the
// flag.  grid: ceil(E * ntile / 256), 256 threads.
__global__ __launch_bounds__(256) void k_tilefar(const int2 *__restrict__ tiles, int ntile, int nblk,
                                                int E, const double *__restrict__ bbox, Box box,
                                                Ewald ew, unsigned char *__restrict__ rfar,
                                                double *__restrict__ epart, int ep_stride,
                                                int ep_off, double *__restrict__ vpart)
{
    const long k = (long)blockIdx.x * 256 + threadIdx.x;
    if (k >= (long)E * ntile) return;
    const int e = (int)(k / ntile), t = (int)(k - (long)e * ntile);
    const int2 ij = tiles[t];
    bool far = false;
    if (ij.x != ij.y) {
        const double *bI = bbox + ((long)e * nblk + ij.x) * RS_BOX;
        const double *bJ = bbox + ((long)e * nblk + ij.y) * RS_BOX;
        double gsq = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double g =
                periodic_gap(bI[d], bI[3 + d], bJ[d], bJ[3 + d], box.L[d], box.inv[d]);
            gsq = fma(g, g, gsq);
        }
        // the margin covers the rounding of the boxes and of rsq
        far = gsq > ew.rcut_sq * (1.0 + 1e-9);
    }
    rfar[k] = far ? 1 : 0;
    if (far) {
        epart[(long)e * ep_stride + ep_off + t] = 0.0;
        if (vpart)
            for (int c = 0; c < 6; ++c) vpart[((long)e * 6 + c) * ep_stride + ep_off + t] = 0.0;
    }
}

template <bool VIR, bool CARG, bool CULL, int W>
__global__ __launch_bounds__(64 * W) void k_realspace(const double *__restrict__ pos, long pstride,
                                                      const double *__restrict__ charge, int n_atoms,
                                                      int NAP, int nblk, double gamma, double half1mg,
                                                      const int2 *__restrict__ tiles, int ntile,
                                                      int t0, int nt,
                                                      Box box, Ewald ew,
                                                      double *__restrict__ rpart,
                                                      double *__restrict__ epart, int ep_stride,
                                                      int ep_off, double *__restrict__ vpart,
                                                      const RsCoef cf,
                                                      const double4 *__restrict__ sites,
                                                      const double *__restrict__ bbox,
                                                      const unsigned char *__restrict__ rfar)
{
    static_assert(W == 2 || W == 4, "two or four waves per tile");
    constexpr int CPW = 64 / W;  // columns per wave
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int e = bid / nt, t = t0 + (bid - e * nt);
    // a tile k_tilefar skipped (its partials zeroed there)
    if (CULL && rfar[(long)e * ntile + t]) return;
    const int I = tiles[t].x, J = tiles[t].y;
    const int l = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const double *X = pos + (long)e * pstride;
    const long base_e = (long)e * nblk * 3 * NAP;
    // CULL: the live 8 x 8 cluster pairs of the wave's columns, from the
    // cluster boxes alone, so their loads overlap the staging of the sites:
    // lane l tests row cluster l & 7 against column cluster l >> 3 (boxes
    // more than r_c apart under every image hold no pair within the cutoff)
    unsigned long long todo = 0ull;
    if constexpr (CULL) {
        const int ci = l & 7, cj = l >> 3;
        const double *cI = bbox + ((long)e * nblk + I) * RS_BOX + 6;
        const double *cJ = bbox + ((long)e * nblk + J) * RS_BOX + 6;
        double gsq = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double g = periodic_gap(cI[8 * d + ci], cI[8 * (3 + d) + ci], cJ[8 * d + cj],
                                          cJ[8 * (3 + d) + cj], box.L[d], box.inv[d]);
            gsq = fma(g, g, gsq);
        }
        const bool live = gsq <= ew.rcut_sq * (1.0 + 1e-9) && I * 64 + 8 * ci < n_atoms &&
                          J * 64 + 8 * cj < n_atoms && (I != J || cj >= ci);
        todo = __ballot(live) & (((CPW == 64 ? 0ull : 1ull << CPW) - 1ull) << (CPW * w));
    }
    __shared__ double4 sI[64], sJ[64];  // x, y, z, charge
    __shared__ double fj[3][64];
    __shared__ double fi[W][3][64];
    __shared__ unsigned short qe[W][64 * CPW];
    __shared__ double ewv[W];
    __shared__ double wv[W][6];
    if (w < 2) {
        const int s = (w == 0 ? J : I) * 64 + l;
        double4 v = make_double4(0.0, 0.0, 0.0, 0.0);
        if (CULL)
            v = sites[(long)e * nblk * 64 + s];
        else if (s < n_atoms)
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
    const int c0 = CPW * w;
    const int i = I * 64 + l;
    int total = 0;
    if constexpr (CULL) {
        // pass 1 over the live cluster pairs: each one's 64 site pairs take
        // one lane each; the queue holds a cluster pair's entries by column,
        // rows within a column
        while (todo) {
            const int bit = __builtin_ctzll(todo);
            todo &= todo - 1ull;
            const int il = 8 * (bit & 7) + (l & 7);
            const int jl = 8 * (bit >> 3) + (l >> 3);
            const double4 a = sI[il], b = sJ[jl];
            const double dx = mimg_fast(a.x - b.x, box.L[0], box.inv[0]);
            const double dy = mimg_fast(a.y - b.y, box.L[1], box.inv[1]);
            const double dz = mimg_fast(a.z - b.z, box.L[2], box.inv[2]);
            const double rsq = fma(dx, dx, fma(dy, dy, dz * dz));
            const int ii = I * 64 + il, j = J * 64 + jl;
            const bool on = j > ii && j < n_atoms && rsq < ew.rcut_sq;
            const unsigned long long bal = __ballot(on);
            if (on) {
                const int r = __builtin_amdgcn_mbcnt_hi(
                    (unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
                qe[w][total + r] = (unsigned short)(il | ((jl - c0) << 6));
            }
            total += __popcll(bal);
        }
    } else {
        // pass 1: in-cutoff columns of row i (j > i; j < n_atoms)
        unsigned int mask = 0u;
        if (i < n_atoms) {
            const double4 a = sI[l];
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
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
        // the wave's pair queue, column-major: column c's rows at off_c + (the
        // lane's rank among the column's set lanes)
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const bool on = (mask >> c) & 1u;
            const unsigned long long bal = __ballot(on);
            if (on) {
                const int r = __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
                qe[w][total + r] = (unsigned short)(l | (c << 6));
            }
            total += __popcll(bal);
        }
    }
    __syncthreads();
    // pass 2: lane l takes queue entries [l*S, l*S + S)
    const bool near = J - I <= 1;
    const int S = (total + 63) >> 6;
    const int q0 = min(l * S, total), q1 = min(q0 + S, total);
    double gx = 0.0, gy = 0.0, gz = 0.0, en = 0.0;
    double vxx = 0.0, vyy = 0.0, vzz = 0.0, vxy = 0.0, vxz = 0.0, vyz = 0.0;
    int jcur = -1;
    double4 b = make_double4(0.0, 0.0, 0.0, 0.0);
    for (int q = q0; q < q1; ++q) {
        const int ent = qe[w][q];
        const int il = ent & 63, jl = c0 + (ent >> 6);
        if (jl != jcur) {
            if (jcur >= 0) {
                atomicAdd(&fj[0][jcur], gx);
                atomicAdd(&fj[1][jcur], gy);
                atomicAdd(&fj[2][jcur], gz);
            }
            gx = gy = gz = 0.0;
            jcur = jl;
            b = sJ[jl];
        }
        const double4 a = sI[il];
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
        gx = fma(-pf, dx, gx);
        gy = fma(-pf, dy, gy);
        gz = fma(-pf, dz, gz);
        if (VIR) {
            vxx = fma(pf * dx, dx, vxx);
            vyy = fma(pf * dy, dy, vyy);
            vzz = fma(pf * dz, dz, vzz);
            vxy = fma(pf * dx, dy, vxy);
            vxz = fma(pf * dx, dz, vxz);
            vyz = fma(pf * dy, dz, vyz);
        }
        atomicAdd(&fi[w][0][il], pf * dx);
        atomicAdd(&fi[w][1][il], pf * dy);
        atomicAdd(&fi[w][2][il], pf * dz);
    }
    if (jcur >= 0) {
        atomicAdd(&fj[0][jcur], gx);
        atomicAdd(&fj[1][jcur], gy);
        atomicAdd(&fj[2][jcur], gz);
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
    if (VIR && threadIdx.x < 6) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < W; ++k) v += wv[k][threadIdx.x];
        vpart[((long)e * 6 + threadIdx.x) * ep_stride + ep_off + t] = v;
    }
    // component d of both slot entries of this tile
    for (int k = threadIdx.x; k < 192; k += 64 * W) {
        const int d = k >> 6, ll = k & 63;
        double si = 0.0;
#pragma unroll
        for (int v = 0; v < W; ++v) si += fi[v][d][ll];
        if (I == J) {
            rpart[base_e + (long)I * 3 * NAP + d * NAP + I * 64 + ll] = si + fj[d][ll];
        } else {
            rpart[base_e + (long)J * 3 * NAP + d * NAP + I * 64 + ll] = si;       // slot J
            rpart[base_e + (long)I * 3 * NAP + d * NAP + J * 64 + ll] = fj[d][ll];  // slot I
        }
    }
    if (threadIdx.x == 0) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < W; ++k) s += ewv[k];
        epart[(long)e * ep_stride + ep_off + t] = s;
    }
}

}  // namespace qtp
