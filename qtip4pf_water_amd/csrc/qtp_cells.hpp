// qtp_cells.hpp -- cell-binned Verlet lists on the device.
//
// The pair test is neighbour.cpp:42-53 bit for bit (minimum image of the
// unwrapped O-O difference, r^2 < r_skin^2 in FP64, no contraction); only the
// candidates differ from the reference's all-pairs loop: the molecules of
// the cells within S cells of the row's cell on a grid of cells at least
// r_skin / S wide (with a relative margin of 1e-9 over the rounding of the
// cell index), which holds every partner within r_skin.  Each row is
// gathered into a bitmap over molecule indices in LDS and compacted from it
// with ballot-free prefix popcounts, so the lists come out in ascending
// order, identical to the all-pairs build.
#pragma once
#include "qtp_common.hpp"

namespace qtp {

// Row bitmaps of the count pass kept for the fill pass up to this size
// (4096 H2O x 32 beads: 64 MiB).
constexpr size_t NL_BM_MAX_BYTES = (size_t)256 << 20;

// Limits of the cell-binned build: the cell counts of one bead in LDS
// (k_cell_build) and four row bitmaps of n_mols bits (k_nl_cell).
constexpr int CELL_MAXCELL = 12288;
constexpr int CELL_MAXMOL = 65536;

struct CellGrid {
    int nc[3];      // cells per dimension (each >= 2 S + 1)
    int S;          // stencil half-width in cells (1 or 2)
    int ncell;      // nc[0] nc[1] nc[2]
    double sc[3];   // nc[d] / L[d]
};

__device__ __forceinline__ int cell_coord(double x, double L, double invL, int nc, double sc)
{
    const double xw = x - L * floor(x * invL);   // [0, L] up to rounding
    int c = (int)(xw * sc);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

// The rebuild test of update.cpp:356-385 (two largest O displacements,
// flag = d1 + d2 > skin; or forced: force_all, or every > 0 and the step
// about to complete is a multiple of every) and, on flagged beads, the
// binning of the beads' O atoms: cell_of, cstart (nc + 1 offsets), the
// molecules of each cell in member (order within a cell arbitrary) and
// their O positions in the same order in mpos (x, y, z, 0), which k_nl_cell
// reads with the member index instead of gathering them from q.
// check = 0: the flags are given in flag[] (no test, nreb untouched).
// grid: P, 1024 threads; dynamic LDS: ncell ints.
__global__ __launch_bounds__(1024) void k_cell_build(int n_mols, int NAP, int NMP,
                                                     const double *__restrict__ q, Box box,
                                                     CellGrid cg, double skin, int check,
                                                     int force_all, int every,
                                                     const unsigned long long *__restrict__ step_ctr,
                                                     const double *__restrict__ disp,
                                                     int *__restrict__ flag,
                                                     double *__restrict__ nreb,
                                                     int *__restrict__ cell_of,
                                                     int *__restrict__ crank,
                                                     int *__restrict__ cstart,
                                                     int *__restrict__ member,
                                                     double4 *__restrict__ mpos)
{
    extern __shared__ int cc[];
    __shared__ double s1[16], s2[16];
    __shared__ int s_flag, s_carry;
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (check) {
        const double *D = disp + (long)b * 3 * NMP;
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
        for (int o = 32; o > 0; o >>= 1) {
            const double b1 = __shfl_xor(m1, o, 64), b2 = __shfl_xor(m2, o, 64);
            top2_merge(m1, m2, b1, b2);
        }
        if (lane == 0) {
            s1[w] = m1;
            s2[w] = m2;
        }
        __syncthreads();
        if (tid == 0) {
            double t1 = s1[0], t2 = s2[0];
            for (int v = 1; v < 16; ++v) top2_merge(t1, t2, s1[v], s2[v]);
            int f = force_all || (sqrt(t1) + sqrt(t2) > skin);
            if (every > 0 && step_ctr && (step_ctr[0] + 1ull) % (unsigned long long)every == 0ull)
                f = 1;
            s_flag = f;
            flag[b] = f;
            if (f && nreb) nreb[b] += 1.0;
        }
    } else if (tid == 0) {
        s_flag = flag[b];
    }
    __syncthreads();
    if (!s_flag) return;
    for (int c = tid; c < cg.ncell; c += 1024) cc[c] = 0;
    __syncthreads();
    const double *Q = q + (long)b * 3 * NAP;
    int *CO = cell_of + (long)b * n_mols, *CR = crank + (long)b * n_mols;
    for (int i = tid; i < n_mols; i += 1024) {
        const int cx = cell_coord(Q[3 * i], box.L[0], box.inv[0], cg.nc[0], cg.sc[0]);
        const int cy = cell_coord(Q[NAP + 3 * i], box.L[1], box.inv[1], cg.nc[1], cg.sc[1]);
        const int cz = cell_coord(Q[2 * NAP + 3 * i], box.L[2], box.inv[2], cg.nc[2], cg.sc[2]);
        const int c = (cz * cg.nc[1] + cy) * cg.nc[0] + cx;
        CO[i] = c;
        CR[i] = atomicAdd(&cc[c], 1);
    }
    if (tid == 0) s_carry = 0;
    __syncthreads();
    // exclusive scan of the counts, 1024 cells at a time (wave scans, then
    // the 16 wave totals by wave 0)
    __shared__ int wt[16];
    int *CS = cstart + (long)b * (cg.ncell + 1);
    for (int c0 = 0; c0 < cg.ncell; c0 += 1024) {
        const int c = c0 + tid;
        const int v = c < cg.ncell ? cc[c] : 0;
        int incl = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane == 63) wt[w] = incl;
        __syncthreads();
        if (w == 0) {
            int x = lane < 16 ? wt[lane] : 0;
            for (int o = 1; o < 16; o <<= 1) {
                const int t = __shfl_up(x, o, 64);
                if (lane >= o) x += t;
            }
            if (lane < 16) wt[lane] = x;   // inclusive over waves
        }
        __syncthreads();
        const int base = s_carry + (w > 0 ? wt[w - 1] : 0);
        if (c < cg.ncell) {
            const int st = base + incl - v;
            CS[c] = st;
            cc[c] = st;
        }
        __syncthreads();
        if (tid == 0) s_carry += wt[15];
        __syncthreads();
    }
    if (tid == 0) CS[cg.ncell] = n_mols;
    __syncthreads();
    int *ME = member + (long)b * n_mols;
    double4 *MP = mpos + (long)b * n_mols;
    for (int i = tid; i < n_mols; i += 1024) {
        const int m = cc[CO[i]] + CR[i];
        ME[m] = i;
        MP[m] = make_double4(Q[3 * i], Q[NAP + 3 * i], Q[2 * NAP + 3 * i], 0.0);
    }
}

// Rows of the lists from the cell bins: one wavefront per molecule row,
// its candidates the (2S+1)^2 runs of 2S+1 x-adjacent cells (contiguous in
// member unless they wrap, then two runs).  FILL = false: row counts
// (cnt[b][0][i] half, j > i; cnt[b][1][i] full, j != i); FILL = true: the
// rows at point / fpoint (k_nl_scan's offsets).  With bmg (P n_mols
// ceil(n_mols/64) words) the count pass also stores each row's bitmap there
// and the fill pass compacts those instead of testing the candidates again.
// grid: P * ceil(n_mols / 4), 256 threads; dynamic LDS: 4 ceil(n_mols/64) u64.
template <bool FILL>
__global__ __launch_bounds__(256) void k_nl_cell(int n_mols, int NAP, const double *__restrict__ q,
                                                 Box box, CellGrid cg, double r_skin_sq,
                                                 int list_size, long fcap,
                                                 const int *__restrict__ flag,
                                                 const int *__restrict__ cell_of,
                                                 const int *__restrict__ cstart,
                                                 const int *__restrict__ member,
                                                 const double4 *__restrict__ mpos,
                                                 int *__restrict__ cnt,
                                                 const int *__restrict__ point,
                                                 const int *__restrict__ fpoint,
                                                 int *__restrict__ list, int *__restrict__ flist,
                                                 unsigned long long *__restrict__ bmg)
{
    extern __shared__ unsigned long long bm_all[];
    __shared__ int rlo[4][64], rpre[4][64];
    const int nrb = (n_mols + 3) / 4;
    const int b = blockIdx.x / nrb;
    if (!flag[b]) return;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = (blockIdx.x - b * nrb) * 4 + w;
    if (i >= n_mols) return;   // wave-local from here on: no block barriers
    const int nw = (n_mols + 63) >> 6;
    unsigned long long *bm = bm_all + (long)w * nw;
    unsigned long long *bmrow = bmg ? bmg + ((long)b * n_mols + i) * nw : nullptr;
    if (!(FILL && bmg)) {
        for (int k = lane; k < nw; k += 64) bm[k] = 0ull;
        const int *CS = cstart + (long)b * (cg.ncell + 1);
        const int *ME = member + (long)b * n_mols;
        const double4 *MP = mpos + (long)b * n_mols;
        const int c = cell_of[(long)b * n_mols + i];
        const int cx = c % cg.nc[0], cy = (c / cg.nc[0]) % cg.nc[1], cz = c / (cg.nc[0] * cg.nc[1]);
        const int span = 2 * cg.S + 1, nrun = 2 * span * span;
        int start = 0, len = 0;
        if (lane < nrun) {
            const int r = lane >> 1, part = lane & 1;
            int yy = cy + r % span - cg.S, zz = cz + r / span - cg.S;
            yy += yy < 0 ? cg.nc[1] : (yy >= cg.nc[1] ? -cg.nc[1] : 0);
            zz += zz < 0 ? cg.nc[2] : (zz >= cg.nc[2] ? -cg.nc[2] : 0);
            const int base = (zz * cg.nc[1] + yy) * cg.nc[0];
            const int x0 = cx - cg.S, x1 = cx + cg.S;
            int lo = 0, hi = -1;
            if (x0 < 0) {
                if (part == 0) { lo = x0 + cg.nc[0]; hi = cg.nc[0] - 1; }
                else { lo = 0; hi = x1; }
            } else if (x1 >= cg.nc[0]) {
                if (part == 0) { lo = x0; hi = cg.nc[0] - 1; }
                else { lo = 0; hi = x1 - cg.nc[0]; }
            } else if (part == 0) {
                lo = x0;
                hi = x1;
            }
            if (hi >= lo) {
                start = CS[base + lo];
                len = CS[base + hi + 1] - start;
            }
        }
        int pre = len;
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(pre, o, 64);
            if (lane >= o) pre += t;
        }
        const int total = __shfl(pre, 63, 64);
        rlo[w][lane] = start;
        rpre[w][lane] = pre - len;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const double *Q = q + (long)b * 3 * NAP;
        const double xi = Q[3 * i], yi = Q[NAP + 3 * i], zi = Q[2 * NAP + 3 * i];
        for (int t0 = 0; t0 < total; t0 += 64) {
            const int t = t0 + lane;
            if (t < total) {
                // last run whose prefix is <= t (runs after it start beyond t)
                int a = 0, e = nrun;
                while (e - a > 1) {
                    const int m = (a + e) >> 1;
                    if (rpre[w][m] <= t) a = m;
                    else e = m;
                }
                const int m = rlo[w][a] + (t - rpre[w][a]);
                const int j = ME[m];
                const double4 pj = MP[m];  // O of j, as in q
                if (j != i) {
                    const double rx = mimg_exact(xi - pj.x, box.L[0]);
                    const double ry = mimg_exact(yi - pj.y, box.L[1]);
                    const double rz = mimg_exact(zi - pj.z, box.L[2]);
                    if (rx * rx + ry * ry + rz * rz < r_skin_sq)
                        atomicOr(&bm[j >> 6], 1ull << (j & 63));
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (bmrow)
            for (int k = lane; k < nw; k += 64) bmrow[k] = bm[k];
    }
    const int wi = i >> 6;
    const unsigned long long above = (i & 63) == 63 ? 0ull : (~0ull << ((i & 63) + 1));
    if (!FILL) {
        int nf = 0, nh = 0;
        for (int k = lane; k < nw; k += 64) {
            const unsigned long long v = bm[k];
            const unsigned long long vh = k > wi ? v : (k == wi ? (v & above) : 0ull);
            nf += __popcll(v);
            nh += __popcll(vh);
        }
        for (int o = 32; o > 0; o >>= 1) {
            nf += __shfl_xor(nf, o, 64);
            nh += __shfl_xor(nh, o, 64);
        }
        if (lane == 0) {
            cnt[((long)b * 2 + 0) * n_mols + i] = nh;
            cnt[((long)b * 2 + 1) * n_mols + i] = nf;
        }
        return;
    }
    int kh = point[(long)b * n_mols + i];
    long kf = fpoint[(long)b * (n_mols + 1) + i];
    int *L = list + (long)b * list_size;
    int *FL = flist + (long)b * fcap;
    for (int k0 = 0; k0 < nw; k0 += 64) {
        const int k = k0 + lane;
        unsigned long long v = k < nw ? (bmrow ? bmrow[k] : bm[k]) : 0ull;
        unsigned long long vh = k > wi ? v : (k == wi ? (v & above) : 0ull);
        const int pf = __popcll(v), ph = __popcll(vh);
        int sf = pf, sh = ph;
        for (int o = 1; o < 64; o <<= 1) {
            const int tf = __shfl_up(sf, o, 64), th = __shfl_up(sh, o, 64);
            if (lane >= o) {
                sf += tf;
                sh += th;
            }
        }
        long of = kf + sf - pf;
        int oh = kh + sh - ph;
        while (v) {
            const int bit = __builtin_ctzll(v);
            v &= v - 1ull;
            if (of < fcap) FL[of] = 64 * k + bit;
            ++of;
        }
        while (vh) {
            const int bit = __builtin_ctzll(vh);
            vh &= vh - 1ull;
            if (oh < list_size) L[oh] = 64 * k + bit;
            ++oh;
        }
        kf += __shfl(sf, 63, 64);
        kh += __shfl(sh, 63, 64);
    }
}

}  // namespace qtp
