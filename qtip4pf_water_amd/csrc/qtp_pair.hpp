// qtp_pair.hpp -- LJ and intramolecular terms, splined pair potentials, energy and
// virial sums.
#pragma once
#include "qtp_kspace.hpp"

namespace qtp {
// LJ over the full O-O list (forces.cpp:489-515, lj_oxygen :25-59) and the
// intramolecular terms (forces.cpp:519-554, :65-138).  LPM lanes per molecule
// share its list row (16, or a whole wave when the launch is small: one
// evaluation of a few hundred molecules leaves most CUs idle otherwise, and
// each lane's walk of the row is a chain of dependent loads); lane 0 of the
// group adds the intramolecular terms.  Each LJ pair is met from both sides;
// its energy is halved.
// grid: E * ceil(n_mols / (256 / LPM)), 256 threads.
constexpr int LJ_BATCH = 4;
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
        // LJ_BATCH entries of the lane's walk per round: their list entries,
        // then their positions, are loaded before any is used (two round
        // trips per round instead of two per entry); the entries are taken
        // in the same order as one at a time
        for (int k = k0 + t; k < k1; k += LJ_BATCH * LPM) {
            int aj[LJ_BATCH];
            double pj[LJ_BATCH][3];
#pragma unroll
            for (int u = 0; u < LJ_BATCH; ++u) {
                const int kk = k + u * LPM;
                aj[u] = kk < k1 ? 3 * fl[kk] : -1;
            }
#pragma unroll
            for (int u = 0; u < LJ_BATCH; ++u)
                for (int d = 0; d < 3; ++d) pj[u][d] = aj[u] >= 0 ? P[d * NAP + aj[u]] : o[d];
#pragma unroll
            for (int u = 0; u < LJ_BATCH; ++u) {
                if (aj[u] < 0) break;
                double r[3];
                for (int d = 0; d < 3; ++d)
                    r[d] = mimg_fast(o[d] - pj[u][d], box.L[d], box.inv[d]);
                const double rsq = fma(r[0], r[0], fma(r[1], r[1], r[2] * r[2]));
                if (rsq < ff.r_cut_sq) {
                    const double rm2 = 1.0 / rsq;
                    const double rm6 = rm2 * rm2 * rm2;
                    const double a12 = rm6 * rm6 * ff.sigmalj12;
                    const double a6 = rm6 * ff.sigmalj6;
                    const double ulj = a12 - a6;
                    elj += ulj * ff.epslj4 - ff.e_cut;
                    const double pf = (ulj + a12) * ff.epslj24 * rm2;
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

}  // namespace qtp
