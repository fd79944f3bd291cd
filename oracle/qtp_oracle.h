/*
 * qtp_oracle.h -- CPU restatement of the qTIP4P/F PIMD hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product path never calls it.
 *
 * Every function follows the arithmetic order of the reference routine it
 * restates (file:line given at each definition in qtp_oracle.c), so that on
 * the same inputs it agrees with the reference to rounding of libm calls.
 *
 * All process globals of the reference (water.cpp:29-38) live in oq_sys.
 * Arrays use the reference layout (indexing.h:6-24):
 *   q[b*n_dof + 9*mol + 3*site + d]  site 0=O, 1=H1, 2=H2
 *   point[b*n_mols + i], list[b*list_size + k], disp[b*3*n_mols + 3*mol + d]
 *
 * Parity status of each function (what pins it) is listed in the header of
 * qtp_oracle.c.
 */
#ifndef QTP_ORACLE_H
#define QTP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    /* sizes */
    int n_mols, n_atoms, n_dof, n_beads, list_size;
    /* box */
    double boxl[3], vol, min_box, max_box;
    /* thermodynamics / integrator (water.cpp:76-83, 143-144) */
    double T_K, beta, beta_n, piovern, omega_p, omega_psq;
    double dt, halfdt, epsilonsc;
    /* ewald inputs (water.cpp:148-157) */
    double eps, acc, rho, kwald_prefactor;
    /* momentum sampling widths (water.cpp:138-139) */
    double sd_ox, sd_h;
    /* 1: force calls add the splined pair potential (get_forces_wspline) */
    int use_spline;
    /* O-O LJ cutoff and Verlet radius (qtip4pfparams.h:41, 106-110): the
     * reference's 9 A and 1.3 of it unless oq_sys_set_lj_rcut changed them */
    double r_cut, r_cut_sq, r_skin, r_skin_sq, e_cut;
    /* Ewald parameters given instead of derived (ew_kappa > 0; not in the
     * reference, whose forces.cpp:574-584 always derives them) */
    double ew_kappa, ew_rcut, ew_kcut;
    int ew_kmax;
} oq_sys;

/* Ewald parameters as derived inside every force call (forces.cpp:574-584). */
typedef struct {
    double kappa, rcut, rcut_sq, k_cut, k_cut_sq, prefactor;
    int k_max;
} oq_ewald;

/* Reference force-field constants (qtip4pfparams.h:43-133). */
typedef struct {
    double pi, angs2bohr, me, kb, Na, eh, kcal2au, mo, mh, K2au, au2fs, fs2au;
    double r_cut, r_cut_sq, r_skin, r_skin_sq;
    double gamma_vs, epslj, sigmalj, sigmalj6, sigmalj12, e_cut;
    double roheq, ar, ar2, ar3, ar4fac, Dr, theta_eq, ktheta, qo, qh, mh2o;
} oq_consts;

void oq_get_consts(oq_consts *c);

/* Cubic/orthorhombic box given directly (the harness form of water.cpp:74-157). */
void oq_sys_init(oq_sys *s, int n_mols, int n_beads, const double boxl[3],
                 double T_K, double dtfs, double eps, double epsilonsc);

void oq_ewald_params(const oq_sys *s, oq_ewald *e);

/* A shorter O-O LJ cutoff (bohr) for boxes narrower than twice the
 * reference's 9 A: r_skin = 1.3 r_cut, e_cut and list_size follow as
 * qtip4pfparams.h:106-110 and water.cpp:135 derive them.  No counterpart in
 * the reference, where the cutoff is a compile-time constant. */
void oq_sys_set_lj_rcut(oq_sys *s, double r_cut);

/* charges[n_atoms] and masses[n_dof] as simulation.cpp:811-822 */
void oq_charges_masses(const oq_sys *s, double *charge, double *mass);

double oq_lj_oxygen(const double r[3], double rsq, double *F, int i, int j);
double oq_intra_morse_bond(const double r[3], double roh, double *F, int i, int bond_step);
double oq_intra_harm_angle(const double r1[3], const double r2[3], double roh1, double roh2,
                           double *F, int i);

/* Splined pair potentials of the fast-CMD path (forces.cpp:452-466,
 * 634-691; tables read by simulation.cpp:36-116).  Each table is the
 * reference's spline file: f[j] = V(a + j h), j = 0..n-1, fitted as Boost.Math's
 * cubic_b_spline does (uniform cubic B-spline interpolation with endpoint
 * derivatives estimated by the 5-point one-sided difference, O(h^4)).
 * oq_spline_set keeps the O-O, O-H and H-H fits for every force call made
 * with sys->use_spline; n >= 5 each.  Returns 0, or -1 on bad sizes. */
int oq_spline_set(const double *oo, int n_oo, const double *oh, int n_oh, const double *hh,
                  int n_hh, double a, double h);
void oq_spline_clear(void);
/* The fit alone: beta[0..n+1] are the B-spline coefficients c_{-1..n}. */
int oq_spline_fit(const double *f, int n, double h, double *beta);
/* Value (and derivative if d != NULL) of a fitted spline at x in [a, a+(n-1)h]. */
double oq_spline_eval(const double *beta, int n, double a, double h, double x, double *d);

/* Single-bead force evaluation (forces.cpp:468-630).  Pointers are the full
 * reference arrays; only the slice of `bead` is read/written. */
double oq_get_forces(const oq_sys *s, const double *q, double *F, const double *charge,
                     const int *point, const int *list, int bead);

/* Components of the single-bead energy, for tests: out[0]=LJ, [1]=intra,
 * [2]=real-space Ewald, [3]=reciprocal, [4]=self, [5]=the total summed with
 * an 80-bit accumulator.  Forces as oq_get_forces. */
double oq_get_forces_parts(const oq_sys *s, const double *q, double *F, const double *charge,
                           const int *point, const int *list, int bead, double out[6]);
/* on != 0: every bead energy (oq_get_forces and the drivers above it) is the
 * 80-bit accumulated total instead of the reference's FP64 running sum. */
void oq_set_exact_energy(int on);

/* Suzuki-Chin force driver (forces.cpp:698-789).  Per-bead reduction is
 * taken in bead order (the reference with OMP_NUM_THREADS=1). */
double oq_force_eval_sc(const oq_sys *s, const double *q, double *Fsc, const double *charge,
                        const double *mass, const int *point, const int *list,
                        double *sc_pot, double *fphys, int nthreads);

/* Verlet list (neighbour.cpp:12-62).  Returns the number of pairs found; if
 * that exceeds list_size the extra pairs are not written (the reference writes
 * them out of bounds) and the return value is still the true count. */
int oq_neighbour_list(const oq_sys *s, const double *q, int *point, int *list,
                      double *disp, int bead);

/* Normal-mode transform (fourier.cpp:12-69), direct DFT in long double. */
void oq_ftransform(const oq_sys *s, double *a, int length, int is_forwards);

/* PILE-G (thermostat.cpp:64-133).  dev = [R, R_add, gamma_draw, xi...],
 * xi ordered bead 1..P-1, dof 0..n_dof-1 (3 + (P-1)*n_dof values). */
double oq_pile_g(const oq_sys *s, double *p, const double *mass, const double *dev);

/* One SC-PIMD step (update.cpp:271-401).  dev as oq_pile_g (ignored if
 * !is_nvt).  Returns the SC conserved potential; *heat -= thermostat return. */
double oq_update_sc(const oq_sys *s, double *q, double *p, const double *mass, double *F,
                    double *fphys, const double *charge, int *point, int *list, double *disp,
                    int is_nvt, const double *dev, double *heat, double *sc_pot,
                    int nthreads, int *n_rebuilds);

/* One classical velocity-Verlet step (update.cpp:78-137). */
double oq_update_class(const oq_sys *s, double *q, double *p, const double *mass, double *F,
                       const double *charge, int *point, int *list, double *disp,
                       int *n_rebuilds);

/* Observables (path_integrals.cpp:29-81). */
double oq_get_kcv(const oq_sys *s, const double *q, const double *f);
double oq_get_conserved(const oq_sys *s, double *q, double *p, const double *mass, double energy);

#ifdef __cplusplus
}
#endif
#endif
