// qtp_step.hpp -- Suzuki-Chin combination, propagation and thermostat, neighbour
// lists, classical step, observables.
#pragma once
#include "qtp_pair.hpp"

namespace qtp {
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
    // this thread's dofs t = threadIdx.x + SC_THREADS r: forces, masses and
    // positions loaded once, before the energy sum (one memory latency)
    constexpr int SC_RUN = 4;
    const int nt = 3 * a.NAP;
    const bool odd = b & 1;
    double fv[SC_RUN], mv[SC_RUN], qv[SC_RUN];
    const bool fits = nt <= SC_RUN * SC_THREADS;
    if (fits) {
#pragma unroll
        for (int r = 0; r < SC_RUN; ++r) {
            const int t = threadIdx.x + r * SC_THREADS;
            if (t < nt) {
                fv[r] = fph[o + t];
                if (odd) {
                    mv[r] = m_atom[t % a.NAP];
                    qv[r] = q[o + t];
                }
            }
        }
    }
    if (threadIdx.x < 64) {
        const double v = eval_energy(epart, b, ep_stride, ep_n, eself);
        if (threadIdx.x == 0) {
            s_V = v;
            Vout[b] = v;
        }
    }
    if (!odd) {
        if (fits) {
#pragma unroll
            for (int r = 0; r < SC_RUN; ++r) {
                const int t = threadIdx.x + r * SC_THREADS;
                if (t < nt) Fsc[o + t] = 2.0 * fv[r] / 3.0;
            }
        } else {
            // larger beads: the same dofs in runs of SC_RUN, each run's loads
            // issued together
            for (int t0 = threadIdx.x; t0 < nt; t0 += SC_RUN * SC_THREADS) {
                double fb[SC_RUN];
#pragma unroll
                for (int r = 0; r < SC_RUN; ++r) {
                    const int t = t0 + r * SC_THREADS;
                    if (t < nt) fb[r] = fph[o + t];
                }
#pragma unroll
                for (int r = 0; r < SC_RUN; ++r) {
                    const int t = t0 + r * SC_THREADS;
                    if (t < nt) Fsc[o + t] = 2.0 * fb[r] / 3.0;
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const double V = s_V;
            scal[b * 4 + 0] = 2.0 * V;
            scal[b * 4 + 1] = V;
            scal[b * 4 + 2] = 0.0;
            scal[b * 4 + 3] = 0.0;
        }
        return;
    }
    // |f|^2 / m and (f / m)^2 over the real atoms: per thread in dof order,
    // then the block sum
    double fsq = 0.0, del = 0.0;
    if (fits) {
#pragma unroll
        for (int r = 0; r < SC_RUN; ++r) {
            const int t = threadIdx.x + r * SC_THREADS;
            if (t < nt && t % a.NAP < a.n_atoms) {
                const double f = fv[r], m = mv[r];
                fsq += f * f / m;
                const double fm = f / m;
                del += fm * fm;
            }
        }
    } else {
        // runs of SC_RUN dofs, loads first, sums in the same dof order
        for (int t0 = threadIdx.x; t0 < nt; t0 += SC_RUN * SC_THREADS) {
            double fb[SC_RUN], mb[SC_RUN];
#pragma unroll
            for (int r = 0; r < SC_RUN; ++r) {
                const int t = t0 + r * SC_THREADS;
                if (t < nt) {
                    fb[r] = fph[o + t];
                    mb[r] = m_atom[t % a.NAP];
                }
            }
#pragma unroll
            for (int r = 0; r < SC_RUN; ++r) {
                const int t = t0 + r * SC_THREADS;
                if (t < nt && t % a.NAP < a.n_atoms) {
                    const double f = fb[r], m = mb[r];
                    fsq += f * f / m;
                    const double fm = f / m;
                    del += fm * fm;
                }
            }
        }
    }
    fsq = block_sum(fsq, sh);
    del = block_sum(del, sh);
    if (threadIdx.x == 0) {
        const double V = s_V;
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
    if (fits) {
#pragma unroll
        for (int r = 0; r < SC_RUN; ++r) {
            const int t = threadIdx.x + r * SC_THREADS;
            if (t < nt) {
                const double dq = sc_scale * fv[r] / mv[r];
                Pp[t] = qv[r] + dq;
                Pm[t] = qv[r] - dq;
            }
        }
    } else {
        for (int t0 = threadIdx.x; t0 < nt; t0 += SC_RUN * SC_THREADS) {
            double fb[SC_RUN], mb[SC_RUN], qb[SC_RUN];
#pragma unroll
            for (int r = 0; r < SC_RUN; ++r) {
                const int t = t0 + r * SC_THREADS;
                if (t < nt) {
                    fb[r] = fph[o + t];
                    mb[r] = m_atom[t % a.NAP];
                    qb[r] = q[o + t];
                }
            }
#pragma unroll
            for (int r = 0; r < SC_RUN; ++r) {
                const int t = t0 + r * SC_THREADS;
                if (t < nt) {
                    const double dq = sc_scale * fb[r] / mb[r];
                    Pp[t] = qb[r] + dq;
                    Pm[t] = qb[r] - dq;
                }
            }
        }
    }
}
// Second half (forces.cpp:776-789): odd beads F_sc = 4f/3 - (f+ - f-) / (9 w^2
// step); the half-step kick p += dt/2 F_sc when requested; an extra last
// block sums the SC potential and sc_pot in bead order (beside the others).
// grid: P * nch + 1, SC_THREADS threads: nch blocks per bead, each taking
// every nch-th run of SC_THREADS dofs (nch > 1 for large systems).
// f2 (may be null): the second half of the split k-space forces of the
// phase-2 evaluations, added to fpm here.
__global__ __launch_bounds__(SC_THREADS) void k_sc2(SCArgs a, const double *__restrict__ fph,
                                                    const double *__restrict__ fpm,
                                                    const double *__restrict__ f2,
                                                    const double *__restrict__ scal,
                                                    double *__restrict__ Fsc, double *__restrict__ p,
                                                    int kick, double *__restrict__ out,
                                                    unsigned long long *__restrict__ step_ctr,
                                                    int nch)
{
    const int b = blockIdx.x / nch, c = blockIdx.x - b * nch;
    const long o = (long)b * 3 * a.NAP;
    const int t0 = c * SC_THREADS + threadIdx.x, dt = nch * SC_THREADS;
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
    // runs of four dofs per thread: the loads of a run are issued together
    const int nt = 3 * a.NAP;
    if (b & 1) {
        const int k = b >> 1;
        const double fscale = scal[b * 4 + 3];
        const long op = (long)(2 * k) * 3 * a.NAP, om = (long)(2 * k + 1) * 3 * a.NAP;
        for (int tr = t0; tr < nt; tr += 4 * dt) {
            double F[4], pv[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int t = tr + r * dt;
                if (t < nt) {
                    F[r] = 4.0 * fph[o + t] / 3.0;
                    const double fp = f2 ? fpm[op + t] + f2[op + t] : fpm[op + t];
                    const double fm = f2 ? fpm[om + t] + f2[om + t] : fpm[om + t];
                    F[r] -= fscale * (fp - fm);
                    if (kick) pv[r] = p[o + t];
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int t = tr + r * dt;
                if (t < nt) {
                    Fsc[o + t] = F[r];
                    if (kick) p[o + t] = pv[r] + a.halfdt * F[r];
                }
            }
        }
    } else if (kick) {
        for (int tr = t0; tr < nt; tr += 4 * dt) {
            double F[4], pv[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int t = tr + r * dt;
                if (t < nt) {
                    F[r] = Fsc[o + t];
                    pv[r] = p[o + t];
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int t = tr + r * dt;
                if (t < nt) p[o + t] = pv[r] + a.halfdt * F[r];
            }
        }
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
    // centroid: cmode 0 free (PIMD / PA-CMD), 2 held fixed (its momentum is
    // zeroed after the kick, no thermostat); cthermo 0 = PILE-G, the Bussi
    // rescale with c = exp(-dt/tau) (c = 0: thermostat.cpp:88-111), 1 =
    // PILE-L, Langevin c1_0 = exp(-dt/tau), c2_0 = sqrt(1 - c1_0^2)
    int cmode, cthermo;
    double cbussi, c1_0, c2_0;
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
                                                 const double2 *__restrict__ mscale,
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
    // propagation mass of mode k: the physical mass, times (w_k / W)^2 for
    // the non-centroid modes of partially adiabatic CMD
    const double m = va ? m_atom[at] * mscale[k].x : 1.0;
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
    if (k == 0) {
        if (A.cmode == 2) up = 0.0;   // centroid held fixed
        cayley_half(uq, up, m, A.halfdt, md);
        const bool pile_g = A.nvt && A.cthermo == 0 && A.cmode != 2;
        double ke0 = (va && pile_g) ? 0.5 * up * up / m : 0.0;
        double chi = 0.0, h = 0.0;
        if (pile_g && !A.host_dev && va) {
            const int ridx = 3 * at + d;
            if (ridx >= 1) {
                const double z = philox_normal(A.seed, *step_ctr, (uint32_t)(P + 1), (uint32_t)ridx);
                chi = z * z;
            }
        }
        if (A.nvt && A.cthermo == 1 && A.cmode != 2 && va) {
            // PILE-L: Langevin on the centroid (random stream 0)
            h = 0.5 * up * up / m;
            const double xi = philox_normal(A.seed, *step_ctr, 0u, (uint32_t)(3 * at + d));
            up = A.c1_0 * up + sqm_atom[at] * A.c2_0 * xi;
            h -= 0.5 * up * up / m;
        }
        nq[f] = uq;
        np[f] = up;
        ke0 = wave_sum(ke0);
        chi = wave_sum(chi);
        h = wave_sum(h);
        if (lane == 0) {
            part[tile] = ke0;
            part[A.ntile + tile] = chi;
            // the centroid's heat slot: PILE-L's here, PILE-G's in k_prop_inv
            if (A.nvt && !pile_g) hpart[tile] = h;
        }
        return;
    }
    cayley_half(uq, up, m, A.halfdt, md);
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
    if (A.nvt && A.cthermo == 0 && A.cmode != 2) {
        const double ke = wave_sum_array(part, A.ntile);
        double R, g2;
        if (A.host_dev) {
            R = dev[0];
            g2 = (A.n_dof % 2 == 0) ? 2.0 * dev[2] + dev[1] * dev[1] : 2.0 * dev[2];
        } else {
            g2 = wave_sum_array(part + A.ntile, A.ntile);
            R = philox_normal(A.seed, *step_ctr, (uint32_t)P, 0u);
        }
        double alpha;
        if (A.cbussi == 0.0) {
            alpha = sqrt((R * R + g2) / (2.0 * A.beta_n * ke));   // thermostat.cpp:99-104
        } else {
            // Bussi, Donadio, Parrinello, J. Chem. Phys. 126, 014101 (2007),
            // eq. (A7): relaxation towards the canonical KE with time constant tau
            const double c = A.cbussi, x = 1.0 / (2.0 * A.beta_n * ke);
            alpha = sqrt(c + (1.0 - c) * (R * R + g2) * x + 2.0 * R * sqrt(c * (1.0 - c) * x));
            if (R + sqrt(c / ((1.0 - c) * x)) < 0.0) alpha = -alpha;
        }
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

// The whole propagation of a step for one 64-dof tile when the centroid needs
// no global rescale (NVE, PILE-L, or the centroid held fixed): kick, forward
// normal-mode transform, sqrt-Cayley half steps around the Langevin steps of
// every mode, inverse transform, displacement tracking (update.cpp:281-352,
// thermostat.cpp:114-129).  The tile's P beads are read from HBM once and
// staged in LDS ([P][64] positions and kicked momenta, then the modes in the
// same space); the arithmetic and its order are k_prop_fwd's and
// k_prop_inv's, so the results are the same bit for bit.
// grid: ntile (XCD-remapped), 64 * PT_WAVES threads, 2 P 64 doubles of LDS;
// P <= PT_WAVES * PT_PER_WAVE.
constexpr int PT_WAVES = 8, PT_PER_WAVE = 8;
// Below this many tiles (one block per CU) the two-kernel form, with a wave
// per (tile, mode), has more parallelism and is faster (216 x 32: 33 tiles,
// 25.6 us fused against 16.4 us for the pair; 4096 x 32: 576 tiles, 69
// against 124 us).
constexpr int PT_MIN_TILES = 256;
inline size_t prop_tile_lds(int P) { return (size_t)2 * P * 64 * sizeof(double); }
__global__ __launch_bounds__(64 * PT_WAVES) void k_prop_tile(
    PropArgs A, const double *__restrict__ Cm, const double4 *__restrict__ mode,
    const double2 *__restrict__ mscale, double *__restrict__ q, double *__restrict__ p,
    const double *__restrict__ F, const double *__restrict__ m_atom,
    const double *__restrict__ sqm_atom, double *__restrict__ part, double *__restrict__ hpart,
    double *__restrict__ disp, int NMP, const unsigned long long *__restrict__ step_ctr)
{
    extern __shared__ double S[];
    const int P = A.P;
    double *SQ = S, *SP = S + (long)P * 64;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long N3 = 3L * A.NAP;
    const int f = tile * 64 + lane;
    const int at = f % A.NAP, d = f / A.NAP;
    const bool va = at < A.n_atoms;
    double qold[PT_PER_WAVE];
#pragma unroll
    for (int j = 0; j < PT_PER_WAVE; ++j) {
        const int b = w + j * PT_WAVES;
        if (b < P) {
            const long idx = b * N3 + f;
            const double qq = q[idx];
            qold[j] = qq;
            SQ[b * 64 + lane] = qq;
            SP[b * 64 + lane] = p[idx] + A.halfdt * F[idx];   // update.cpp:281-287
        }
    }
    __syncthreads();
    double rq[PT_PER_WAVE], rp[PT_PER_WAVE];
#pragma unroll
    for (int j = 0; j < PT_PER_WAVE; ++j) {
        const int k = w + j * PT_WAVES;
        if (k >= P) continue;
        const double m = va ? m_atom[at] * mscale[k].x : 1.0;
        const double *cr = Cm + (long)k * P;
        double uq = 0.0, up = 0.0;
        for (int b = 0; b < P; ++b) {
            const double c = cr[b];
            uq = fma(c, SQ[b * 64 + lane], uq);
            up = fma(c, SP[b * 64 + lane], up);
        }
        const double4 md = mode[k];
        if (k == 0) {
            if (A.cmode == 2) up = 0.0;
            cayley_half(uq, up, m, A.halfdt, md);
            double h = 0.0;
            if (A.nvt && A.cthermo == 1 && A.cmode != 2 && va) {
                h = 0.5 * up * up / m;
                const double xi = philox_normal(A.seed, *step_ctr, 0u, (uint32_t)(3 * at + d));
                up = A.c1_0 * up + sqm_atom[at] * A.c2_0 * xi;
                h -= 0.5 * up * up / m;
            }
            h = wave_sum(h);
            if (lane == 0) {
                part[tile] = 0.0;
                part[A.ntile + tile] = 0.0;
                if (A.nvt) hpart[tile] = h;
            }
            cayley_half(uq, up, va ? m_atom[at] : 1.0, A.halfdt, mode[0]);
        } else {
            cayley_half(uq, up, m, A.halfdt, md);
            if (A.nvt) {
                double h = 0.0;
                if (va) {
                    h += 0.5 * up * up / m;
                    const double xi =
                        philox_normal(A.seed, *step_ctr, (uint32_t)k, (uint32_t)(3 * at + d));
                    up = md.z * up + sqm_atom[at] * md.w * xi;
                    h -= 0.5 * up * up / m;
                }
                h = wave_sum(h);
                if (lane == 0) hpart[(long)k * A.ntile + tile] = h;
            }
            cayley_half(uq, up, m, A.halfdt, md);
        }
        rq[j] = uq;
        rp[j] = up;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PT_PER_WAVE; ++j) {
        const int k = w + j * PT_WAVES;
        if (k < P) {
            SQ[k * 64 + lane] = rq[j];
            SP[k * 64 + lane] = rp[j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PT_PER_WAVE; ++j) {
        const int b = w + j * PT_WAVES;
        if (b >= P) continue;
        double xq = Cm[b] * SQ[lane], xp = Cm[b] * SP[lane];
        for (int k = 1; k < P; ++k) {
            const double c = Cm[(long)k * P + b];
            xq = fma(c, SQ[k * 64 + lane], xq);
            xp = fma(c, SP[k * 64 + lane], xp);
        }
        const long idx = b * N3 + f;
        q[idx] = xq;
        p[idx] = xp;
        if (va && (at % 3) == 0) disp[((long)b * 3 + d) * NMP + at / 3] += xq - qold[j];
    }
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
// grid: ceil(len/64), 256 threads, nm_generic_lds(P) bytes of LDS (the
// block's 64 columns of all P beads; the matrix entries are wave-uniform
// loads from global memory).  fwd: u_k = sum_b C[k][b] x_b, else the transpose.
inline size_t nm_generic_lds(int P) { return (size_t)P * 64 * sizeof(double); }
__global__ __launch_bounds__(256) void k_nm_generic(int P, int len, const double *__restrict__ Cm,
                                                    double *__restrict__ a, int fwd)
{
    extern __shared__ double X[];
    const int lane = threadIdx.x & 63, kg = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + lane;
    for (int b = kg; b < P; b += 4) X[b * 64 + lane] = c < len ? a[(long)b * len + c] : 0.0;
    __syncthreads();
    for (int k = kg; k < P; k += 4) {
        double u = 0.0;
        for (int b = 0; b < P; ++b) {
            const double cf = fwd ? Cm[k * P + b] : Cm[b * P + k];
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

// Per bead: two largest O displacement norms; flag = d1 + d2 > r_skin - r_cut,
// or forced (force_all; or every > 0 and the step about to complete, step_ctr
// + 1, a multiple of every).
// grid: P, 256 threads.
__global__ __launch_bounds__(256) void k_check(int n_mols, int NMP, const double *__restrict__ disp,
                                               double skin, int force_all, int every,
                                               const unsigned long long *__restrict__ step_ctr,
                                               int *__restrict__ flag, double *__restrict__ nreb)
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
        int f = force_all || (sqrt(s1[0]) + sqrt(s2[0]) > skin);
        if (every > 0 && step_ctr && (step_ctr[0] + 1ull) % (unsigned long long)every == 0ull)
            f = 1;
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
                                                     int every,
                                                     const unsigned long long *__restrict__ step_ctr,
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
        int f = force_all || (sqrt(t1) + sqrt(t2) > skin);
        if (every > 0 && step_ctr && (step_ctr[0] + 1ull) % (unsigned long long)every == 0ull)
            f = 1;
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
// With a thermostat (lang != 0) a Langevin step p = c1 p + sqrt(m/beta) c2 xi
// comes first (random stream 0 of the step counter; the counter advances in
// k_vv_b, so every block reads the same value here).
struct VVLang {
    int on;
    double c1, c2;
    uint64_t seed;
};
__global__ __launch_bounds__(256) void k_vv_a(int n_atoms, int NAP, int NMP, double halfdt,
                                              double dt, const double *__restrict__ m_atom,
                                              const double *__restrict__ F, double *__restrict__ q,
                                              double *__restrict__ p, double *__restrict__ disp,
                                              VVLang lang, const double *__restrict__ sqm_atom,
                                              const unsigned long long *__restrict__ step_ctr)
{
    const int f = blockIdx.x * 256 + threadIdx.x;
    if (f >= 3 * NAP) return;
    const int at = f % NAP, d = f / NAP;
    if (at >= n_atoms) return;
    double p0 = p[f];
    if (lang.on) {
        const double xi = philox_normal(lang.seed, *step_ctr, 0u, (uint32_t)(3 * at + d));
        p0 = lang.c1 * p0 + sqm_atom[at] * lang.c2 * xi;
    }
    double pp = p0 + halfdt * F[f];
    p[f] = pp;
    const double m = m_atom[at];
    q[f] += pp * dt / m;
    if (at % 3 == 0) disp[d * NMP + at / 3] += pp * dt / m;
}

// F = f_phys of the single evaluation, p += dt/2 F, energy out.
// fx, ex (may be null): forces (planar) and n_ex energy partials of a
// further term added to V (the learned correction's), summed in a fixed order.
__global__ __launch_bounds__(256) void k_vv_b(int n_atoms, int NAP, double halfdt,
                                              const double *__restrict__ fe,
                                              const double *__restrict__ epart, int ep_stride,
                                              int ep_n, double eself,
                                              double *__restrict__ F, double *__restrict__ p,
                                              double *__restrict__ out,
                                              unsigned long long *__restrict__ step_ctr,
                                              const double *__restrict__ fx,
                                              const float *__restrict__ ex, int n_ex)
{
    const int f = blockIdx.x * 256 + threadIdx.x;
    if (f < 3 * NAP) {
        const double v = fx ? fe[f] + fx[f] : fe[f];
        F[f] = v;
        const int at = f % NAP;
        if (at < n_atoms) p[f] += halfdt * v;
    }
    if (blockIdx.x == 0 && threadIdx.x < 64) {
        double v = eval_energy(epart, 0, ep_stride, ep_n, eself);
        if (ex) {
            double x = 0.0;
            for (int a = threadIdx.x; a < n_ex; a += 64) x += (double)ex[a];
            v += wave_sum(x);
        }
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

// ---------------------------------------------------------------------------
// Centroid mean force, radial distribution functions (no counterpart in the
// reference; the CMD and g(r) tools its README leaves to other programs)
// ---------------------------------------------------------------------------
// acc[t] += (1/P) sum_b F[b][t]: the force on the centroid of dof t, summed
// over the steps of a run (the mean force of centroid-constrained runs).
// grid: ceil(3 NAP / 256), 256 threads.
__global__ __launch_bounds__(256) void k_cmf_acc(int P, int NAP, const double *__restrict__ F,
                                                 double *__restrict__ acc,
                                                 unsigned long long *__restrict__ n)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) n[0] += 1ull;
    if (t >= 3 * NAP) return;
    const long N3 = 3L * NAP;
    double s = 0.0;
    for (int b = 0; b < P; ++b) s += F[b * N3 + t];
    acc[t] += s / P;
}

// Centroid positions (bead average) of the planar state.  grid: ceil(3 NAP/256).
__global__ __launch_bounds__(256) void k_centroid(int P, int NAP, const double *__restrict__ q,
                                                  double *__restrict__ qc)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= 3 * NAP) return;
    const long N3 = 3L * NAP;
    double s = 0.0;
    for (int b = 0; b < P; ++b) s += q[b * N3 + t];
    qc[t] = s / P;
}

// Pair-distance histograms of the centroid positions: O-O over molecule
// pairs i < j, O-H over every (O of i, H of j) including j = i.  Bin
// floor(r / dr) of the minimum-image distance (bohr) below n_bins * dr.
// One wavefront per molecule i (lanes over j); integer counts, so the
// result does not depend on the order of the atomic adds.
// grid: ceil(n_mols / 4), 256 threads.
__global__ __launch_bounds__(256) void k_rdf(int n_mols, int NAP, const double *__restrict__ qc,
                                             Box box, double inv_dr, int n_bins,
                                             unsigned int *__restrict__ hoo,
                                             unsigned int *__restrict__ hoh)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n_mols) return;
    const double xo = qc[3 * i], yo = qc[NAP + 3 * i], zo = qc[2 * NAP + 3 * i];
    for (int j = lane; j < n_mols; j += 64) {
        for (int s = 0; s < 3; ++s) {
            if (s == 0 && j <= i) continue;
            const int a = 3 * j + s;
            const double dx = mimg_exact(qc[a] - xo, box.L[0]);
            const double dy = mimg_exact(qc[NAP + a] - yo, box.L[1]);
            const double dz = mimg_exact(qc[2 * NAP + a] - zo, box.L[2]);
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            const double bf = r * inv_dr;
            if (bf < (double)n_bins) atomicAdd(s == 0 ? &hoo[(int)bf] : &hoh[(int)bf], 1u);
        }
    }
}

}  // namespace qtp
