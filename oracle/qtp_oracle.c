/*
 * qtp_oracle.c -- CPU restatement of the qTIP4P/F PIMD hot path (test oracle).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product library.
 *
 * Each routine follows the operation order of the reference routine cited at
 * its definition, so results agree with the reference up to libm rounding.
 * Build with -ffp-contract=off (no FMA contraction), as the reference is built
 * for plain x86-64.
 *
 * What pins each part (tests/ holds the checks):
 *   constants / derived globals  : published qTIP4P/F parameters, water.cpp
 *                                  formulas evaluated independently in Python
 *   LJ, bond, angle, M-site      : closed forms (LJ zero and minimum, Morse
 *                                  Taylor series, angle at theta_eq), central
 *                                  finite differences, sum F = 0
 *   real + reciprocal Ewald      : brute-force numpy sum over explicit k-vectors
 *                                  and explicit images (long double), kappa
 *                                  independence of the converged total
 *   SC driver                    : finite-difference gradient of the SC
 *                                  potential, even/odd weights
 *   neighbour list               : compiled reference neighbour.cpp (bit-exact)
 *   ftransform                   : orthonormality, exact inverse, explicit
 *                                  cos/sin matrix entries
 *   pile_g                       : compiled reference thermostat.cpp, fed the
 *                                  same mt19937 stream (bit-exact)
 *   update_sc / update_class     : exact 2x2 Cayley map per mode, free ring
 *                                  polymer, NVE drift
 *   get_kcv / get_conserved      : algebraic definitions (parity pinned by
 *                                  closed forms only)
 */
#define _GNU_SOURCE
#include "qtp_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------------
 * Constants: qtip4pfparams.h:41-133.  Evaluated once, in the same order.
 * ------------------------------------------------------------------------- */
static oq_consts C;
static int C_ready = 0;

static void init_consts(void)
{
    if (C_ready) return;
    C.pi = 3.14159265359;                       /* truncated on purpose, :45 */
    C.angs2bohr = 1.0 / 0.529177210903;
    C.me = 9.10938356e-31;
    C.kb = 1.38064852e-23;
    C.Na = 6.0221409e23;
    C.eh = 4.3597447222071e-18;
    C.kcal2au = 1.5936e-3;
    const double mhgmol = 1.0080, mogmol = 15.9994;
    C.mo = 0.001 * mogmol / (C.Na * C.me);
    C.mh = 0.001 * mhgmol / (C.Na * C.me);
    C.K2au = C.kb / C.eh;
    const double hbar_SI = 1.054571817e-34;
    C.au2fs = hbar_SI * 1e15 / C.eh;
    C.fs2au = 1.0 / C.au2fs;
    C.mh2o = (mhgmol * 2.0 + mogmol) / C.Na;
    C.r_cut = 9.0 * C.angs2bohr;
    C.r_cut_sq = C.r_cut * C.r_cut;
    C.r_skin = 1.30 * C.r_cut;
    C.r_skin_sq = C.r_skin * C.r_skin;
    C.gamma_vs = 0.73612;
    C.epslj = 0.1852 * C.kcal2au;
    C.sigmalj = 3.1589 * C.angs2bohr;
    C.sigmalj6 = pow(C.sigmalj, 6.0);
    C.sigmalj12 = C.sigmalj6 * C.sigmalj6;
    C.e_cut = 4.0 * C.epslj * (pow((C.sigmalj / C.r_cut), 12.0) - pow((C.sigmalj / C.r_cut), 6.0));
    C.roheq = 0.9419 * C.angs2bohr;
    C.ar = 2.287 / C.angs2bohr;
    C.ar2 = C.ar * C.ar;
    C.ar3 = C.ar2 * C.ar;
    C.ar4fac = 7.0 * C.ar3 * C.ar / 12.0;
    C.Dr = 116.09 * C.kcal2au;
    C.theta_eq = 107.4 * C.pi / 180.0;
    C.ktheta = 87.85 * C.kcal2au;
    C.qo = -1.1128;
    C.qh = -0.5 * C.qo;
    C_ready = 1;
}

void oq_get_consts(oq_consts *c)
{
    init_consts();
    *c = C;
}

/* water.cpp:74-157 with the box given directly (cubic or orthorhombic). */
void oq_sys_init(oq_sys *s, int n_mols, int n_beads, const double boxl[3],
                 double T_K, double dtfs, double eps, double epsilonsc)
{
    init_consts();
    memset(s, 0, sizeof(*s));
    s->n_mols = n_mols;
    s->n_atoms = 3 * n_mols;
    s->n_dof = 3 * s->n_atoms;
    s->n_beads = n_beads;
    s->boxl[0] = boxl[0];
    s->boxl[1] = boxl[1];
    s->boxl[2] = boxl[2];
    s->vol = boxl[0] * boxl[1] * boxl[2];
    double rho_ox = n_mols / s->vol;
    s->list_size = (int)round(4.0 * C.r_skin * C.r_skin * C.r_skin * n_mols * rho_ox);
    s->r_cut = C.r_cut;
    s->r_cut_sq = C.r_cut_sq;
    s->r_skin = C.r_skin;
    s->r_skin_sq = C.r_skin_sq;
    s->e_cut = C.e_cut;
    s->T_K = T_K;
    double T_red = C.K2au * T_K;
    s->beta = 1.0 / T_red;
    s->piovern = C.pi / n_beads;
    s->beta_n = s->beta / n_beads;
    s->omega_p = 1.0 / (s->beta_n * 1.0);
    s->omega_psq = s->omega_p * s->omega_p;
    s->sd_ox = sqrt(C.mo / s->beta_n);
    s->sd_h = sqrt(C.mh / s->beta_n);
    s->dt = dtfs * C.fs2au;
    s->halfdt = 0.5 * s->dt;
    s->epsilonsc = epsilonsc;
    s->eps = eps;
    s->acc = sqrt(-log(eps));
    s->rho = s->n_atoms / s->vol;
    double min_xy = fmin(boxl[0], boxl[1]);
    s->min_box = fmin(min_xy, boxl[2]);
    s->kwald_prefactor = 2.0 * C.pi / s->vol;
    double max_xy = fmax(boxl[0], boxl[1]);
    s->max_box = fmax(max_xy, boxl[2]);
}

/* qtip4pfparams.h:106-110 and water.cpp:135 with the cutoff given */
void oq_sys_set_lj_rcut(oq_sys *s, double r_cut)
{
    init_consts();
    s->r_cut = r_cut;
    s->r_cut_sq = r_cut * r_cut;
    s->r_skin = 1.30 * r_cut;
    s->r_skin_sq = s->r_skin * s->r_skin;
    s->e_cut = 4.0 * C.epslj * (pow((C.sigmalj / r_cut), 12.0) - pow((C.sigmalj / r_cut), 6.0));
    double rho_ox = s->n_mols / s->vol;
    s->list_size = (int)round(4.0 * s->r_skin * s->r_skin * s->r_skin * s->n_mols * rho_ox);
}

/* forces.cpp:574-584 */
void oq_ewald_params(const oq_sys *s, oq_ewald *e)
{
    init_consts();
    if (s->ew_kappa > 0.0) {
        e->kappa = s->ew_kappa;
        e->rcut = s->ew_rcut;
        e->rcut_sq = s->ew_rcut * s->ew_rcut;
        e->k_cut = s->ew_kcut;
        e->k_cut_sq = s->ew_kcut * s->ew_kcut;
        e->k_max = s->ew_kmax;
        e->prefactor = 2.0 * s->ew_kappa / sqrt(C.pi);
        return;
    }
    double kappa = 2.6 * pow(s->rho / s->vol, 1.0 / 6.0);
    double rcut = s->acc / kappa;
    if (rcut > 0.5 * s->min_box) {
        rcut = 0.5 * s->min_box;
        kappa = s->acc / rcut;
    }
    e->kappa = kappa;
    e->rcut = rcut;
    e->rcut_sq = rcut * rcut;
    e->k_cut = 2.0 * s->acc * kappa;
    e->k_cut_sq = e->k_cut * e->k_cut;
    e->k_max = (int)(e->k_cut * s->max_box / (2.0 * C.pi));
    e->prefactor = 2.0 * kappa / sqrt(C.pi);
}

/* simulation.cpp:811-822 */
void oq_charges_masses(const oq_sys *s, double *charge, double *mass)
{
    init_consts();
    for (int i = 0; i < s->n_mols; ++i) {
        charge[3 * i] = C.qo;
        charge[3 * i + 1] = C.qh;
        charge[3 * i + 2] = C.qh;
        for (int d = 0; d < 3; ++d) {
            mass[9 * i + d] = C.mo;
            mass[9 * i + d + 3] = C.mh;
            mass[9 * i + d + 6] = C.mh;
        }
    }
}

static inline void min_image(const oq_sys *s, double r[3])
{
    for (int d = 0; d < 3; ++d) r[d] -= s->boxl[d] * round(r[d] / s->boxl[d]);
}

/* ---------------------------------------------------------------------------
 * Pair and intramolecular terms.  F points at the bead slice.
 * ------------------------------------------------------------------------- */

/* forces.cpp:25-59 */
static double lj_oxygen_cut(const double r[3], double rsq, double *F, int i, int j, double e_cut);

double oq_lj_oxygen(const double r[3], double rsq, double *F, int i, int j)
{
    init_consts();
    return lj_oxygen_cut(r, rsq, F, i, j, C.e_cut);
}

static double lj_oxygen_cut(const double r[3], double rsq, double *F, int i, int j, double e_cut)
{
    double rm2 = 1.0 / rsq;
    double rm6 = rm2 * rm2 * rm2;
    double rm12 = rm6 * rm6;
    rm6 *= C.sigmalj6;
    rm12 *= C.sigmalj12;
    double energy = rm12 - rm6;
    double pf = energy + rm12;
    energy *= 4.0 * C.epslj;
    energy -= e_cut;
    pf *= 24.0 * C.epslj * rm2;
    for (int d = 0; d < 3; ++d) {
        F[9 * i + d] += pf * r[d];
        F[9 * j + d] -= pf * r[d];
    }
    return energy;
}

/* forces.cpp:65-93 */
double oq_intra_morse_bond(const double r[3], double roh, double *F, int i, int bond_step)
{
    init_consts();
    double rm1 = 1.0 / roh;
    double x = roh - C.roheq;
    double x2 = x * x;
    double x3 = x2 * x;
    double x4 = x2 * x2;
    double energy = C.ar2 * x2 - C.ar3 * x3 + C.ar4fac * x4;
    energy *= C.Dr;
    double pf = 2.0 * C.ar2 * x - 3.0 * C.ar3 * x2 + 4.0 * C.ar4fac * x3;
    pf *= C.Dr * rm1;
    for (int d = 0; d < 3; ++d) {
        F[9 * i + d] -= pf * r[d];
        F[9 * i + 3 * bond_step + d] += pf * r[d];
    }
    return energy;
}

/* forces.cpp:96-138 */
double oq_intra_harm_angle(const double r1[3], const double r2[3], double roh1, double roh2,
                           double *F, int i)
{
    init_consts();
    double dot = r1[0] * r2[0] + r1[1] * r2[1] + r1[2] * r2[2];
    double r1m1 = 1.0 / roh1;
    double r2m1 = 1.0 / roh2;
    double r1m2 = r1m1 * r1m1;
    double r2m2 = r2m1 * r2m1;
    double r12m1 = r1m1 * r2m1;
    double c = dot * r12m1;
    double c2 = c * c;
    double dxdc = 1.0 / sqrt(1.0 - c2);
    double c2r1 = dot * r1m2 * r12m1;
    double c2r2 = dot * r2m2 * r12m1;
    double theta = acos(c);
    double td = theta - C.theta_eq;
    double energy = 0.5 * C.ktheta * td * td;
    double pf = C.ktheta * td * dxdc;
    for (int d = 0; d < 3; ++d) {
        F[9 * i + d] += pf * (r12m1 * (r1[d] + r2[d]) - c2r1 * r1[d] - c2r2 * r2[d]);
        F[9 * i + d + 3] -= pf * (r12m1 * r2[d] - c2r1 * r1[d]);
        F[9 * i + d + 6] -= pf * (r12m1 * r1[d] - c2r2 * r2[d]);
    }
    return energy;
}

/* forces.cpp:141-177.  ef is the site-force slice (3 per site). */
static double rwald(const double r[3], double rsq, double *ef, const double *charge,
                    int i, int j, double kappa)
{
    double kappasq = kappa * kappa;
    double pref = 2.0 * kappa / sqrt(C.pi);
    double rij = sqrt(rsq);
    double rm1 = 1.0 / rij;
    double rm2 = rm1 * rm1;
    double qiqj = charge[i] * charge[j];
    double erfc_val = erfc(kappa * rij);
    if (i / 3 == j / 3) erfc_val -= 1.0;
    double energy = qiqj * rm1 * erfc_val;
    double pf = energy + qiqj * pref * exp(-kappasq * rsq);
    pf *= rm2;
    for (int d = 0; d < 3; ++d) {
        ef[3 * i + d] += pf * r[d];
        ef[3 * j + d] -= pf * r[d];
    }
    return energy;
}

/* forces.cpp:180-448.  pos/ef are bead slices of site positions/forces. */
static double kwald(const oq_sys *s, const double *pos, double *ef, const double *charge,
                    int k_max, double kappa, double k_cut_sq)
{
    const int n_atoms = s->n_atoms, n_dof = s->n_dof, K = k_max + 1;
    double kappasq = kappa * kappa;
    double q4 = 0.25 / kappasq;
    double energy = 0.0;
    double bk[3], kr[3], ck[3], sk[3], rk[3];
    double *cosk = (double *)malloc(sizeof(double) * n_dof * K);
    double *sink = (double *)malloc(sizeof(double) * n_dof * K);
    double (*xy)[4] = malloc(sizeof(double[4]) * n_atoms);
    double (*xyz)[8] = malloc(sizeof(double[8]) * n_atoms);
#define CK(i, j) cosk[(i) * K + (j)]
#define SK(i, j) sink[(i) * K + (j)]
    for (int d = 0; d < 3; ++d) bk[d] = 2.0 * C.pi / s->boxl[d];

    for (int i = 0; i < n_atoms; ++i) {
        for (int d = 0; d < 3; ++d) {
            CK(3 * i + d, 0) = 1.0;
            SK(3 * i + d, 0) = 0.0;
        }
        for (int d = 0; d < 3; ++d) kr[d] = bk[d] * pos[3 * i + d];
        for (int d = 0; d < 3; ++d) {
            ck[d] = cos(kr[d]);
            sk[d] = sin(kr[d]);
        }
        for (int j = 1; j < K; ++j) {
            for (int d = 0; d < 3; ++d) {
                CK(3 * i + d, j) = ck[d] * CK(3 * i + d, j - 1) - sk[d] * SK(3 * i + d, j - 1);
                SK(3 * i + d, j) = sk[d] * CK(3 * i + d, j - 1) + ck[d] * SK(3 * i + d, j - 1);
            }
        }
        for (int j = 0; j < K; ++j) {
            CK(3 * i, j) *= charge[i];
            SK(3 * i, j) *= charge[i];
        }
    }

    const double kp = s->kwald_prefactor;
    for (int kx = 0; kx < K; ++kx) {
        rk[0] = bk[0] * kx;
        double ksq = rk[0] * rk[0];
        /* kx only */
        if (ksq > 0.0 && ksq < k_cut_sq) {
            double sr = 0.0, si = 0.0;
            double ef0 = exp(-ksq * q4) / ksq;
            if (kx != 0) ef0 *= 2.0;
            for (int i = 0; i < n_atoms; ++i) {
                sr += CK(3 * i, kx);
                si += SK(3 * i, kx);
            }
            energy += kp * ef0 * (sr * sr + si * si);
            for (int i = 0; i < n_atoms; ++i) {
                double pf = (SK(3 * i, kx) * sr - CK(3 * i, kx) * si) * 2.0 * kp * ef0 * rk[0];
                ef[3 * i] += pf;
            }
        }
        /* ky = 0, kz > 0 */
        for (int kz = 1; kz < K; ++kz) {
            rk[2] = bk[2] * kz;
            ksq = rk[0] * rk[0] + rk[2] * rk[2];
            if (ksq > k_cut_sq) break;
            double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
            for (int i = 0; i < n_atoms; ++i) {
                double cc = CK(3 * i, kx) * CK(3 * i + 2, kz);
                double cs = CK(3 * i, kx) * SK(3 * i + 2, kz);
                double sc = SK(3 * i, kx) * CK(3 * i + 2, kz);
                double ss = SK(3 * i, kx) * SK(3 * i + 2, kz);
                xy[i][0] = cc - ss;
                xy[i][1] = cs + sc;
                xy[i][2] = cc + ss;
                xy[i][3] = sc - cs;
                s1 += xy[i][0];
                s2 += xy[i][1];
                s3 += xy[i][2];
                s4 += xy[i][3];
            }
            double ef1 = kp * exp(-ksq * q4) / ksq;
            if (kx != 0) ef1 *= 2.0;
            energy += ef1 * (s1 * s1 + s2 * s2 + s3 * s3 + s4 * s4);
            double xt = 2.0 * ef1 * rk[0];
            double zt = 2.0 * ef1 * rk[2];
            for (int i = 0; i < n_atoms; ++i) {
                double t1 = s2 * xy[i][0] - s1 * xy[i][1];
                double t2 = s4 * xy[i][2] - s3 * xy[i][3];
                ef[3 * i] -= xt * (t1 + t2);
                ef[3 * i + 2] -= zt * (t1 - t2);
            }
        }
        /* ky > 0 */
        for (int ky = 1; ky < K; ++ky) {
            rk[1] = bk[1] * ky;
            ksq = rk[0] * rk[0] + rk[1] * rk[1];
            if (ksq > k_cut_sq) break;
            double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
            for (int i = 0; i < n_atoms; ++i) {
                double cc = CK(3 * i, kx) * CK(3 * i + 1, ky);
                double cs = CK(3 * i, kx) * SK(3 * i + 1, ky);
                double sc = SK(3 * i, kx) * CK(3 * i + 1, ky);
                double ss = SK(3 * i, kx) * SK(3 * i + 1, ky);
                xy[i][0] = cc - ss;
                xy[i][1] = cs + sc;
                xy[i][2] = cc + ss;
                xy[i][3] = sc - cs;
                s1 += xy[i][0];
                s2 += xy[i][1];
                s3 += xy[i][2];
                s4 += xy[i][3];
            }
            double ef2 = kp * exp(-ksq * q4) / ksq;
            if (kx != 0) ef2 *= 2.0;
            energy += ef2 * (s1 * s1 + s2 * s2 + s3 * s3 + s4 * s4);
            double xt = 2.0 * ef2 * rk[0];
            double yt = 2.0 * ef2 * rk[1];
            for (int i = 0; i < n_atoms; ++i) {
                double t1 = s2 * xy[i][0] - s1 * xy[i][1];
                double t2 = s4 * xy[i][2] - s3 * xy[i][3];
                ef[3 * i] -= xt * (t1 + t2);
                ef[3 * i + 1] -= yt * (t1 - t2);
            }
            /* ky > 0, kz > 0 */
            for (int kz = 1; kz < K; ++kz) {
                rk[2] = bk[2] * kz;
                ksq = rk[0] * rk[0] + rk[1] * rk[1] + rk[2] * rk[2];
                if (ksq > k_cut_sq) break;
                double u[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                for (int i = 0; i < n_atoms; ++i) {
                    double cz = CK(3 * i + 2, kz), sz = SK(3 * i + 2, kz);
                    double cca = xy[i][0] * cz;
                    double csa = xy[i][0] * sz;
                    double sca = xy[i][1] * cz;
                    double ssa = xy[i][1] * sz;
                    double ccb = xy[i][2] * cz;
                    double csb = xy[i][2] * sz;
                    double scb = xy[i][3] * cz;
                    double ssb = xy[i][3] * sz;
                    xyz[i][0] = cca - ssa;
                    xyz[i][1] = sca + csa;
                    xyz[i][2] = cca + ssa;
                    xyz[i][3] = sca - csa;
                    xyz[i][4] = ccb - ssb;
                    xyz[i][5] = scb + csb;
                    xyz[i][6] = ccb + ssb;
                    xyz[i][7] = scb - csb;
                    for (int m = 0; m < 8; ++m) u[m] += xyz[i][m];
                }
                double ef3 = kp * exp(-ksq * q4) / ksq;
                if (kx != 0) ef3 *= 2.0;
                energy += ef3 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2] + u[3] * u[3] +
                                 u[4] * u[4] + u[5] * u[5] + u[6] * u[6] + u[7] * u[7]);
                double xt = 2.0 * ef3 * rk[0];
                double yt = 2.0 * ef3 * rk[1];
                double zt = 2.0 * ef3 * rk[2];
                for (int i = 0; i < n_atoms; ++i) {
                    double t1 = u[1] * xyz[i][0] - u[0] * xyz[i][1];
                    double t2 = u[3] * xyz[i][2] - u[2] * xyz[i][3];
                    double t3 = u[5] * xyz[i][4] - u[4] * xyz[i][5];
                    double t4 = u[7] * xyz[i][6] - u[6] * xyz[i][7];
                    ef[3 * i] -= xt * (t1 + t2 + t3 + t4);
                    ef[3 * i + 1] -= yt * (t1 + t2 - t3 - t4);
                    ef[3 * i + 2] -= zt * (t1 - t2 + t3 - t4);
                }
            }
        }
    }
#undef CK
#undef SK
    free(cosk);
    free(sink);
    free(xy);
    free(xyz);
    return energy;
}

/* Energies in extended precision (test aid, not in the reference): when set,
 * oq_get_forces returns the bead energy summed with an 80-bit accumulator in
 * the same order, so that its rounding error (~1e-19 relative per term) lies
 * far below the FP64 running sum's (the reference's, ~1e-16 per term, which
 * over the millions of real-space pair terms of a 4096-molecule box reaches
 * ~1e-11 Eh). */
static int oq_exact = 0;
void oq_set_exact_energy(int on) { oq_exact = on; }

/* forces.cpp:468-630 */
double oq_get_forces_parts(const oq_sys *s, const double *q_all, double *F_all, const double *charge,
                           const int *point_all, const int *list_all, int bead, double out[6])
{
    init_consts();
    const int n_mols = s->n_mols, n_atoms = s->n_atoms, n_dof = s->n_dof;
    const double *q = q_all + (size_t)bead * n_dof;
    double *F = F_all + (size_t)bead * n_dof;
    const int *point = point_all + (size_t)bead * n_mols;
    const int *list = list_all + (size_t)bead * s->list_size;
    double *cp = (double *)malloc(sizeof(double) * n_dof);
    double *ef = (double *)malloc(sizeof(double) * n_dof);
    double r[3], r1[3], r2[3];
    double energy = 0.0;
    long double ex = 0.0L;
    double parts[6] = {0, 0, 0, 0, 0, 0};

    for (int i = 0; i < n_dof; ++i) F[i] = 0.0;

    /* LJ over the half list, forces.cpp:489-515 */
    for (int i = 0; i < n_mols - 1; ++i) {
        for (int nb = point[i]; nb < point[i + 1]; ++nb) {
            int j = list[nb];
            for (int d = 0; d < 3; ++d) r[d] = q[9 * i + d] - q[9 * j + d];
            min_image(s, r);
            double rsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            if (rsq < s->r_cut_sq) {
                double e = lj_oxygen_cut(r, rsq, F, i, j, s->e_cut);
                energy += e;
                ex += e;
                parts[0] += e;
            }
        }
    }

    /* intramolecular, forces.cpp:519-554 */
    for (int i = 0; i < n_mols; ++i) {
        for (int d = 0; d < 3; ++d) r1[d] = q[9 * i + d] - q[9 * i + d + 3];
        min_image(s, r1);
        double r1sq = r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2];
        double roh1 = sqrt(r1sq);
        double e = oq_intra_morse_bond(r1, roh1, F, i, 1);
        energy += e;
        ex += e;
        parts[1] += e;
        for (int d = 0; d < 3; ++d) r2[d] = q[9 * i + d] - q[9 * i + d + 6];
        min_image(s, r2);
        double r2sq = r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2];
        double roh2 = sqrt(r2sq);
        e = oq_intra_morse_bond(r2, roh2, F, i, 2);
        energy += e;
        ex += e;
        parts[1] += e;
        e = oq_intra_harm_angle(r1, r2, roh1, roh2, F, i);
        energy += e;
        ex += e;
        parts[1] += e;
    }

    /* M-site positions, forces.cpp:558-564 */
    for (int i = 0; i < n_mols; ++i) {
        for (int d = 0; d < 3; ++d) {
            cp[9 * i + d] = C.gamma_vs * q[9 * i + d] +
                            0.5 * (1.0 - C.gamma_vs) * (q[9 * i + d + 3] + q[9 * i + d + 6]);
            cp[9 * i + d + 3] = q[9 * i + d + 3];
            cp[9 * i + d + 6] = q[9 * i + d + 6];
        }
    }
    for (int i = 0; i < n_dof; ++i) ef[i] = 0.0;

    oq_ewald ew;
    oq_ewald_params(s, &ew);

    /* real space, all site pairs, forces.cpp:589-608 */
    for (int i = 0; i < n_atoms - 1; ++i) {
        for (int j = i + 1; j < n_atoms; ++j) {
            for (int d = 0; d < 3; ++d) r[d] = cp[3 * i + d] - cp[3 * j + d];
            min_image(s, r);
            double rsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            if (rsq < ew.rcut_sq) {
                double e = rwald(r, rsq, ef, charge, i, j, ew.kappa);
                energy += e;
                ex += e;
                parts[2] += e;
            }
        }
    }

    /* reciprocal space, forces.cpp:611 */
    double ek = kwald(s, cp, ef, charge, ew.k_max, ew.kappa, ew.k_cut_sq);
    energy += ek;
    ex += ek;
    parts[3] = ek;

    /* chain rule, forces.cpp:613-619 */
    for (int i = 0; i < n_mols; ++i) {
        for (int d = 0; d < 3; ++d) {
            F[9 * i + d] += C.gamma_vs * ef[9 * i + d];
            F[9 * i + d + 3] += ef[9 * i + d + 3] + 0.5 * (1.0 - C.gamma_vs) * ef[9 * i + d];
            F[9 * i + d + 6] += ef[9 * i + d + 6] + 0.5 * (1.0 - C.gamma_vs) * ef[9 * i + d];
        }
    }

    /* self term, forces.cpp:622-624 */
    for (int i = 0; i < n_atoms; ++i) {
        double e = 0.5 * ew.prefactor * charge[i] * charge[i];
        energy -= e;
        ex -= e;
        parts[4] -= e;
    }

    free(cp);
    free(ef);
    parts[5] = (double)ex;
    if (out) memcpy(out, parts, sizeof(parts));
    return oq_exact ? (double)ex : energy;
}

/* ---------------------------------------------------------------------------
 * Splined pair potentials (forces.cpp:452-466, 634-691).
 * ------------------------------------------------------------------------- */
static struct {
    int ready;
    int n[3];
    double a, h, cutsq[3];
    double *beta[3];
} SPL;

/* Uniform cubic B-spline interpolation of f[0..n-1] at x_j = a + j h with
 * clamped ends, s'(a) and s'(b) from the 5-point one-sided differences
 * (-25 f0 + 48 f1 - 36 f2 + 16 f3 - 3 f4) / 12h and its mirror: the
 * published construction of Boost.Math's cubic_b_spline (no version is
 * pinned by the reference).  Interpolation rows c_{j-1} + 4 c_j + c_{j+1} =
 * 6 f_j with c_{-1} = c_1 - 2h s'(a) and c_n = c_{n-2} + 2h s'(b); the
 * tridiagonal system is solved in long double. */
int oq_spline_fit(const double *f, int n, double h, double *beta)
{
    if (n < 5 || !(h > 0.0)) return -1;
    const long double hl = h;
    const long double a1 = (-25.0L * f[0] + 48.0L * f[1] - 36.0L * f[2] + 16.0L * f[3] -
                            3.0L * f[4]) / (12.0L * hl);
    const long double b1 = (25.0L * f[n - 1] - 48.0L * f[n - 2] + 36.0L * f[n - 3] -
                            16.0L * f[n - 4] + 3.0L * f[n - 5]) / (12.0L * hl);
    long double *dg = (long double *)malloc(sizeof(long double) * n);
    long double *up = (long double *)malloc(sizeof(long double) * n);
    long double *lo = (long double *)malloc(sizeof(long double) * n);
    long double *rhs = (long double *)malloc(sizeof(long double) * n);
    for (int j = 0; j < n; ++j) {
        dg[j] = 4.0L;
        up[j] = 1.0L;
        lo[j] = 1.0L;
        rhs[j] = 6.0L * f[j];
    }
    up[0] = 2.0L;
    rhs[0] += 2.0L * hl * a1;
    lo[n - 1] = 2.0L;
    rhs[n - 1] -= 2.0L * hl * b1;
    /* Thomas algorithm */
    for (int j = 1; j < n; ++j) {
        const long double m = lo[j] / dg[j - 1];
        dg[j] -= m * up[j - 1];
        rhs[j] -= m * rhs[j - 1];
    }
    long double *c = (long double *)malloc(sizeof(long double) * (n + 2));
    c[n] = rhs[n - 1] / dg[n - 1];
    for (int j = n - 2; j >= 0; --j) c[j + 1] = (rhs[j] - up[j] * c[j + 2]) / dg[j];
    c[0] = c[2] - 2.0L * hl * a1;
    c[n + 1] = c[n - 1] + 2.0L * hl * b1;
    for (int j = 0; j < n + 2; ++j) beta[j] = (double)c[j];
    free(dg);
    free(up);
    free(lo);
    free(rhs);
    free(c);
    return 0;
}

double oq_spline_eval(const double *beta, int n, double a, double h, double x, double *d)
{
    double t = (x - a) / h;
    int k = (int)floor(t);
    if (k > n - 2) k = n - 2;  /* x = b uses the last interval */
    if (k < 0) k = 0;
    const double u = t - k;
    const double v = 1.0 - u;
    /* cubic B-spline weights of c_{k-1}, c_k, c_{k+1}, c_{k+2} */
    const double w0 = v * v * v / 6.0;
    const double w1 = (3.0 * u * u * u - 6.0 * u * u + 4.0) / 6.0;
    const double w2 = (-3.0 * u * u * u + 3.0 * u * u + 3.0 * u + 1.0) / 6.0;
    const double w3 = u * u * u / 6.0;
    const double *c = beta + k; /* c[0] = c_{k-1} */
    if (d) {
        const double d0 = -0.5 * v * v;
        const double d1 = 1.5 * u * u - 2.0 * u;
        const double d2 = -1.5 * u * u + u + 0.5;
        const double d3 = 0.5 * u * u;
        *d = (c[0] * d0 + c[1] * d1 + c[2] * d2 + c[3] * d3) / h;
    }
    return c[0] * w0 + c[1] * w1 + c[2] * w2 + c[3] * w3;
}

void oq_spline_clear(void)
{
    for (int t = 0; t < 3; ++t) {
        free(SPL.beta[t]);
        SPL.beta[t] = NULL;
    }
    SPL.ready = 0;
}

int oq_spline_set(const double *oo, int n_oo, const double *oh, int n_oh, const double *hh,
                  int n_hh, double a, double h)
{
    const double *tab[3] = {oo, oh, hh};
    const int nn[3] = {n_oo, n_oh, n_hh};
    oq_spline_clear();
    for (int t = 0; t < 3; ++t) {
        if (nn[t] < 5) return -1;
        SPL.beta[t] = (double *)malloc(sizeof(double) * (nn[t] + 2));
        if (oq_spline_fit(tab[t], nn[t], h, SPL.beta[t]) != 0) return -1;
        SPL.n[t] = nn[t];
        /* simulation.cpp:57-58: cut = start + step (n - 1), squared */
        const double cut = a + h * (nn[t] - 1);
        SPL.cutsq[t] = cut * cut;
    }
    SPL.a = a;
    SPL.h = h;
    SPL.ready = 1;
    return 0;
}

/* forces.cpp:643-689 then get_spline :452-466 for each pair in range */
static double spline_pairs(const oq_sys *s, const double *q, double *F)
{
    double energy = 0.0, r[3];
    for (int i = 0; i < s->n_atoms; ++i) {
        for (int j = i + 1; j < s->n_atoms; ++j) {
            for (int d = 0; d < 3; ++d) r[d] = q[3 * i + d] - q[3 * j + d];
            min_image(s, r);
            const double rsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            const int t = (i % 3 == 0) ? ((j % 3 == 0) ? 0 : 1) : ((j % 3 == 0) ? 1 : 2);
            if (rsq < SPL.cutsq[t]) {
                const double rij = sqrt(rsq);
                double fp;
                energy += oq_spline_eval(SPL.beta[t], SPL.n[t], SPL.a, SPL.h, rij, &fp);
                fp *= 1.0 / rij;
                for (int d = 0; d < 3; ++d) {
                    F[3 * i + d] -= fp * r[d];
                    F[3 * j + d] += fp * r[d];
                }
            }
        }
    }
    return energy;
}

double oq_get_forces(const oq_sys *s, const double *q, double *F, const double *charge,
                     const int *point, const int *list, int bead)
{
    double e = oq_get_forces_parts(s, q, F, charge, point, list, bead, NULL);
    if (s->use_spline && SPL.ready)
        e += spline_pairs(s, q + (size_t)bead * s->n_dof, F + (size_t)bead * s->n_dof);
    return e;
}

/* forces.cpp:698-789.  The reference reduces tot_energy and vop over an OMP
 * team; here each bead's terms are stored and summed in bead order, which is
 * the reference's result with one thread. */
double oq_force_eval_sc(const oq_sys *s, const double *q, double *Fsc, const double *charge,
                        const double *mass, const int *point, const int *list,
                        double *sc_pot, double *fphys, int nthreads)
{
    const int P = s->n_beads, n_dof = s->n_dof;
    double *qp = (double *)malloc(sizeof(double) * (size_t)P * n_dof);
    double *qm = (double *)malloc(sizeof(double) * (size_t)P * n_dof);
    double *fp = (double *)malloc(sizeof(double) * (size_t)P * n_dof);
    double *fm = (double *)malloc(sizeof(double) * (size_t)P * n_dof);
    double *tot_b = (double *)calloc(P, sizeof(double));
    double *vop_b = (double *)calloc(P, sizeof(double));
    double *fsq_b = (double *)calloc(P, sizeof(double));
    int b;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (b = 0; b < P; ++b) {
        const size_t o = (size_t)b * n_dof;
        double energy = oq_get_forces(s, q, fphys, charge, point, list, b);
        if (b % 2 == 0) {
            tot_b[b] = 2.0 * energy;
            for (int i = 0; i < n_dof; ++i) Fsc[o + i] = 2.0 * fphys[o + i] / 3.0;
            vop_b[b] = energy;
        } else {
            double t = 4.0 * energy;
            double fsq = 0.0;
            for (int i = 0; i < n_dof; ++i) fsq += fphys[o + i] * fphys[o + i] / mass[i];
            fsq /= 3.0 * s->omega_psq;
            /* the reference adds 4V and fsq to the reduction variable in two
             * separate statements: keep that association */
            tot_b[b] = t;
            fsq_b[b] = fsq;
            double delta = 0.0;
            for (int i = 0; i < n_dof; ++i) {
                double dt_ = fphys[o + i] / mass[i];
                delta += dt_ * dt_;
            }
            delta /= s->n_atoms * P;
            delta = 1.0 / sqrt(delta);
            double sc_scale = s->epsilonsc * delta;
            double fscale = 1.0 / (9.0 * s->omega_psq * sc_scale);
            for (int i = 0; i < n_dof; ++i) {
                double dq = sc_scale * fphys[o + i] / mass[i];
                qp[o + i] = q[o + i] + dq;
                qm[o + i] = q[o + i] - dq;
            }
            oq_get_forces(s, qp, fp, charge, point, list, b);
            oq_get_forces(s, qm, fm, charge, point, list, b);
            for (int i = 0; i < n_dof; ++i) {
                Fsc[o + i] = 4.0 * fphys[o + i] / 3.0;
                Fsc[o + i] -= fscale * (fp[o + i] - fm[o + i]);
            }
        }
    }
    double tot = 0.0, vop = 0.0;
    for (b = 0; b < P; ++b) {
        if (b % 2 == 0) {
            tot += tot_b[b];
            vop += vop_b[b];
        } else {
            tot += tot_b[b];
            tot += fsq_b[b];
        }
    }
    tot /= 3.0;
    *sc_pot = 2.0 * vop / P;
    free(qp);
    free(qm);
    free(fp);
    free(fm);
    free(tot_b);
    free(vop_b);
    free(fsq_b);
    return tot;
}

/* neighbour.cpp:12-62 */
int oq_neighbour_list(const oq_sys *s, const double *q_all, int *point_all, int *list_all,
                      double *disp_all, int bead)
{
    init_consts();
    const int n_mols = s->n_mols, ls = s->list_size;
    const double *q = q_all + (size_t)bead * s->n_dof;
    int *point = point_all + (size_t)bead * n_mols;
    int *list = list_all + (size_t)bead * ls;
    double *disp = disp_all + (size_t)bead * 3 * n_mols;
    double r[3];
    for (int i = 0; i < ls; ++i) list[i] = 0;
    for (int i = 0; i < 3 * n_mols; ++i) disp[i] = 0.0;
    int k = 0;
    for (int i = 0; i < n_mols; ++i) {
        point[i] = k;
        for (int j = i + 1; j < n_mols; ++j) {
            for (int d = 0; d < 3; ++d) r[d] = q[9 * i + d] - q[9 * j + d];
            min_image(s, r);
            double rsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            if (rsq < s->r_skin_sq) {
                if (k < ls) list[k] = j;
                k += 1;
            }
        }
    }
    return k;
}

/* fourier.cpp:12-69.  FFTW's R2HC/HC2R (documented halfcomplex convention:
 * hc[k] = Re X_k for k <= P/2, hc[P-k] = Im X_k, X_k = sum_j x_j e^{-2 pi i jk/P};
 * HC2R is the unnormalised inverse) evaluated as a direct DFT in long double
 * with exact pi, then scaled exactly as the reference does. */
void oq_ftransform(const oq_sys *s, double *a, int length, int is_forwards)
{
    const int P = s->n_beads;
    double scale = sqrt(1.0 / P);
    double twoscale = sqrt(2.0 / P);
    double root2 = sqrt(2.0);
    double *x = (double *)malloc(sizeof(double) * P);
    double *hc = (double *)malloc(sizeof(double) * P);
    long double *cs = (long double *)malloc(sizeof(long double) * P);
    long double *sn = (long double *)malloc(sizeof(long double) * P);
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int m = 0; m < P; ++m) {
        cs[m] = cosl(twopi * m / P);
        sn[m] = sinl(twopi * m / P);
    }
    for (int i = 0; i < length; ++i) {
        for (int b = 0; b < P; ++b) x[b] = a[(size_t)b * length + i];
        if (is_forwards) {
            for (int k = 0; k <= P / 2; ++k) {
                long double re = 0.0L, im = 0.0L;
                for (int j = 0; j < P; ++j) {
                    int m = (int)(((long)j * k) % P);
                    re += (long double)x[j] * cs[m];
                    im -= (long double)x[j] * sn[m];
                }
                hc[k] = (double)re;
                if (k > 0 && k < P - k) hc[P - k] = (double)im;
            }
            a[i] = scale * hc[0];
            for (int b = 1; b < P; ++b) a[(size_t)b * length + i] = twoscale * hc[b];
            if (P % 2 == 0) a[(size_t)P * length / 2 + i] /= root2;
        } else {
            for (int b = 0; b < P; ++b) hc[b] = x[b];
            for (int b = 1; b < P; ++b) hc[b] /= root2;
            if (P % 2 == 0) hc[P / 2] *= root2;
            for (int j = 0; j < P; ++j) {
                long double v = hc[0];
                for (int k = 1; k < P - k; ++k) {
                    int m = (int)(((long)j * k) % P);
                    v += 2.0L * ((long double)hc[k] * cs[m] - (long double)hc[P - k] * sn[m]);
                }
                if (P % 2 == 0) v += (long double)hc[P / 2] * ((j % 2) ? -1.0L : 1.0L);
                x[j] = (double)v;
            }
            for (int b = 0; b < P; ++b) a[(size_t)b * length + i] = scale * x[b];
        }
    }
    free(x);
    free(hc);
    free(cs);
    free(sn);
}

/* thermostat.cpp:64-133, with the random deviates supplied by the caller in
 * the order the reference draws them. */
double oq_pile_g(const oq_sys *s, double *p, const double *mass, const double *dev)
{
    init_consts();
    const int n_atoms = s->n_atoms, n_dof = s->n_dof, P = s->n_beads;
    double prefactor = 4.0 / (s->beta_n * 1.0);
    double heat = 0.0;
    double ke = 0.0;
    for (int i = 0; i < n_atoms; ++i)
        for (int d = 0; d < 3; ++d) ke += 0.5 * p[3 * i + d] * p[3 * i + d] / mass[3 * i + d];
    heat += ke;
    double R = dev[0];
    int dof = 3 * n_atoms;
    double g2;
    if (dof % 2 == 0) {
        double R_add = dev[1];
        g2 = 2.0 * dev[2] + R_add * R_add;
    } else {
        g2 = 2.0 * dev[2];
    }
    double alpha_sq = (R * R + g2) / (2.0 * s->beta_n * ke);
    double alpha = sqrt(alpha_sq);
    for (int i = 0; i < n_dof; ++i) p[i] *= alpha;
    for (int i = 0; i < n_atoms; ++i)
        for (int d = 0; d < 3; ++d) heat -= 0.5 * p[3 * i + d] * p[3 * i + d] / mass[3 * i];
    const double *xi = dev + 3;
    for (int b = 1; b < P; ++b) {
        double *pb = p + (size_t)b * n_dof;
        for (int i = 0; i < n_dof; ++i) heat += 0.5 * pb[i] * pb[i] / mass[i];
        double gamma_k = prefactor * sin(b * C.pi / P);
        double c1 = exp(-s->dt * gamma_k);
        double c2 = sqrt(1.0 - c1 * c1);
        for (int i = 0; i < n_dof; ++i) {
            double sq = sqrt(mass[i] / s->beta_n);
            pb[i] = c1 * pb[i] + sq * c2 * xi[(size_t)(b - 1) * n_dof + i];
        }
        for (int i = 0; i < n_dof; ++i) heat -= 0.5 * pb[i] * pb[i] / mass[i];
    }
    return heat;
}

/* update.cpp:301-320 / :329-348, one sqrt-Cayley half step in NM space */
static void cayley_half(const oq_sys *s, double *q, double *p, const double *mass, double *disp)
{
    const int P = s->n_beads, n_atoms = s->n_atoms, n_dof = s->n_dof, n_mols = s->n_mols;
    double prefactor = 2.0 / (1.0 * s->beta_n);
    double prefactor_sq = prefactor * prefactor;
    for (int b = 0; b < P; ++b) {
        double sn = sin(b * s->piovern);
        double wsq = prefactor_sq * sn * sn;
        double scale = 1.0 / sqrt(1.0 + 0.25 * wsq * s->dt * s->dt);
        for (int i = 0; i < n_atoms; ++i) {
            for (int d = 0; d < 3; ++d) {
                size_t o = (size_t)b * n_dof + 3 * i + d;
                double temp = s->halfdt * p[o] / mass[3 * i] + q[o];
                temp *= scale;
                p[o] = p[o] - s->halfdt * wsq * q[o] * mass[3 * i];
                p[o] *= scale;
                if (i % 3 == 0) disp[(size_t)b * 3 * n_mols + i + d] += temp - q[o];
                q[o] = temp;
            }
        }
    }
}

/* update.cpp:356-385 (SC) and :97-123 (classical): top-2 displacement test */
static int needs_rebuild(const oq_sys *s, const double *disp)
{
    init_consts();
    double m1 = 0.0, m2 = 0.0;
    for (int i = 0; i < s->n_mols; ++i) {
        double d2 = disp[3 * i] * disp[3 * i] + disp[3 * i + 1] * disp[3 * i + 1] +
                    disp[3 * i + 2] * disp[3 * i + 2];
        if (d2 > m1) {
            m2 = m1;
            m1 = d2;
        } else if (d2 > m2) {
            m2 = d2;
        }
    }
    m1 = sqrt(m1);
    m2 = sqrt(m2);
    return (m1 + m2 > s->r_skin - s->r_cut);
}

/* update.cpp:271-401 */
double oq_update_sc(const oq_sys *s, double *q, double *p, const double *mass, double *F,
                    double *fphys, const double *charge, int *point, int *list, double *disp,
                    int is_nvt, const double *dev, double *heat, double *sc_pot,
                    int nthreads, int *n_rebuilds)
{
    const int P = s->n_beads, n_dof = s->n_dof, n_mols = s->n_mols;
    const size_t N = (size_t)P * n_dof;
    for (size_t o = 0; o < N; ++o) p[o] += s->halfdt * F[o];
    oq_ftransform(s, q, n_dof, 1);
    oq_ftransform(s, p, n_dof, 1);
    oq_ftransform(s, disp, 3 * n_mols, 1);
    cayley_half(s, q, p, mass, disp);
    if (is_nvt) *heat -= oq_pile_g(s, p, mass, dev);
    cayley_half(s, q, p, mass, disp);
    oq_ftransform(s, q, n_dof, 0);
    oq_ftransform(s, p, n_dof, 0);
    oq_ftransform(s, disp, 3 * n_mols, 0);
    for (int b = 0; b < P; ++b) {
        if (needs_rebuild(s, disp + (size_t)b * 3 * n_mols)) {
            oq_neighbour_list(s, q, point, list, disp, b);
            if (n_rebuilds) (*n_rebuilds)++;
        }
    }
    double energy = oq_force_eval_sc(s, q, F, charge, mass, point, list, sc_pot, fphys, nthreads);
    for (size_t o = 0; o < N; ++o) p[o] += s->halfdt * F[o];
    return energy;
}

/* update.cpp:78-137 */
double oq_update_class(const oq_sys *s, double *q, double *p, const double *mass, double *F,
                       const double *charge, int *point, int *list, double *disp,
                       int *n_rebuilds)
{
    const int n_atoms = s->n_atoms;
    for (int i = 0; i < n_atoms; ++i) {
        for (int d = 0; d < 3; ++d) {
            p[3 * i + d] += s->halfdt * F[3 * i + d];
            q[3 * i + d] += p[3 * i + d] * s->dt / mass[3 * i + d];
            if (i % 3 == 0) disp[i + d] += p[3 * i + d] * s->dt / mass[3 * i + d];
        }
    }
    if (needs_rebuild(s, disp)) {
        oq_neighbour_list(s, q, point, list, disp, 0);
        if (n_rebuilds) (*n_rebuilds)++;
    }
    double energy = oq_get_forces(s, q, F, charge, point, list, 0);
    for (int i = 0; i < 3 * n_atoms; ++i) p[i] += s->halfdt * F[i];
    return energy;
}

/* path_integrals.cpp:10-48 */
double oq_get_kcv(const oq_sys *s, const double *q, const double *f)
{
    const int P = s->n_beads, n_dof = s->n_dof;
    double *qc = (double *)malloc(sizeof(double) * n_dof);
    for (int i = 0; i < n_dof; ++i) {
        double v = 0.0;
        for (int b = 0; b < P; ++b) v += q[(size_t)b * n_dof + i];
        v /= (double)(float)P; /* static_cast<float>(n_beads), path_integrals.cpp:18 */
        qc[i] = v;
    }
    double tq = 0.0;
    for (int b = 0; b < P; b += 2) {
        for (int j = 0; j < n_dof; ++j) {
            double dq = q[(size_t)b * n_dof + j] - qc[j];
            tq -= dq * f[(size_t)b * n_dof + j];
        }
    }
    tq /= P;
    free(qc);
    return 0.5 * n_dof / s->beta + tq;
}

/* path_integrals.cpp:52-81 (round-trips q and p through the NM transform) */
double oq_get_conserved(const oq_sys *s, double *q, double *p, const double *mass, double energy)
{
    init_consts();
    const int P = s->n_beads, n_dof = s->n_dof;
    double piovern = C.pi / P;
    oq_ftransform(s, q, n_dof, 1);
    oq_ftransform(s, p, n_dof, 1);
    double nm_kin = 0.0, v_spring = 0.0;
    for (int b = 0; b < P; b++) {
        double q2 = 0.0;
        double pref = 2.0 / (1.0 * s->beta_n);
        double sn = sin(b * piovern);
        double wsq = pref * pref * sn * sn;
        for (int j = 0; j < n_dof; ++j) {
            size_t o = (size_t)b * n_dof + j;
            nm_kin += 0.5 * p[o] * p[o] / mass[j];
            q2 += mass[j] * q[o] * q[o];
        }
        q2 *= 0.5 * wsq;
        v_spring += q2;
    }
    oq_ftransform(s, q, n_dof, 0);
    oq_ftransform(s, p, n_dof, 0);
    return energy + nm_kin + v_spring;
}
