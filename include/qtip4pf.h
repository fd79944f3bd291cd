/*
 * qtip4pf.h -- C ABI of the MI355X (gfx950) qTIP4P/F path-integral step.
 *
 * A drop-in for the hot-path calls of the rayseashyn/qtip4pf_water C++ driver
 * (simulation.cpp / water.cpp stay the caller).  Every reference routine the
 * driver calls has an entry point here taking the same arrays in the same
 * layout (indexing.h:6-24, bead-major, 9 doubles per molecule O,H1,H2):
 *
 *   q[b*n_dof + 9*mol + 3*site + d], point[b*n_mols + i],
 *   list[b*list_size + k], disp[b*3*n_mols + 3*mol + d]
 *
 * The reference keeps its parameters in process globals (water.cpp:29-38);
 * here they live in a context created from qtp_params.  All arithmetic is
 * FP64.  Functions return QTP_OK (0) or a negative status; the message of the
 * last failure is available from qtp_last_error().  No exceptions cross the
 * ABI.  A context serves one host thread at a time.
 *
 * Two ways to drive it:
 *   - reference-shaped calls with host arrays (qtp_get_forces_nospline ...),
 *     each a full host->device->host round trip, for the reference driver;
 *   - resident mode (qtp_set_state / qtp_run / qtp_get_state): state stays
 *     in HBM, the whole step is one replayed HIP graph, no per-step host sync.
 */
#ifndef QTIP4PF_H
#define QTIP4PF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QTP_OK 0
#define QTP_ERR_ARG (-1)           /* bad argument or size */
#define QTP_ERR_HIP (-2)           /* HIP runtime failure */
#define QTP_ERR_LIST_OVERFLOW (-3) /* Verlet list larger than list_size */
#define QTP_ERR_NONFINITE (-4)     /* NaN/Inf in positions or forces */
#define QTP_ERR_STATE (-5)         /* call out of order (e.g. no state set) */

typedef struct qtp_ctx qtp_ctx;

/* Inputs of water.cpp:44-66 that the step depends on, plus the box (the
 * reference derives it from its lattice, water.cpp:90-120; here it is given).
 * Zero/negative optional fields take the reference's derivation. */
typedef struct {
    int n_mols;
    int n_beads;          /* 1 = classical, even >= 2 = Suzuki-Chin (water.cpp:69-72) */
    double boxl[3];       /* bohr */
    double T_K;           /* config.txt T_K */
    double dtfs;          /* config.txt dtfs */
    double eps_ewald;     /* config.txt eps_ewald (water.cpp:148) */
    double epsilon_sc;    /* config.txt epsilon_sc */
    int list_size;        /* 0: round(4 r_skin^3 n_mols rho_ox) (water.cpp:135) */
    /* Ewald override; ewald_kappa <= 0 derives (kappa, r_c, k_cut, k_max)
     * inside each force call as forces.cpp:574-584 does.  Either way
     * 1 <= k_max <= 45 (the k-space factor tables live in LDS), r_c <= L/2
     * and kappa * r_c <= 6.2; qtp_create rejects other values. */
    double ewald_kappa, ewald_rcut, ewald_kcut;
    int ewald_kmax;
    /* O-O Lennard-Jones cutoff (bohr); <= 0: the reference's compile-time
     * 9 A (qtip4pfparams.h:41).  A shorter one keeps boxes narrower than 18 A
     * physical (the reference only warns, water.cpp:176-178); the Verlet
     * radius is 1.3 of it and e_cut and the list size follow it as
     * qtip4pfparams.h:106-110 and water.cpp:135 derive them. */
    double lj_rcut;
} qtp_params;

/* Derived globals (water.cpp:74-157) and Ewald parameters (forces.cpp:574-584). */
typedef struct {
    int n_mols, n_atoms, n_dof, n_beads, list_size;
    double boxl[3], vol, min_box, max_box;
    double beta, beta_n, piovern, omega_p, omega_psq, dt, halfdt, epsilonsc;
    double acc, rho, kwald_prefactor, sd_ox, sd_h;
    double kappa, rcut, k_cut;
    int k_max, n_kgroups;
    double lj_rcut, r_skin;   /* LJ cutoff and Verlet radius (bohr) */
} qtp_derived;

int qtp_create(const qtp_params *prm, qtp_ctx **out);
void qtp_destroy(qtp_ctx *ctx);
/* ctx == NULL: the message of the last failed context-free call
 * (qtp_get_kcv_noalloc) on the calling thread. */
const char *qtp_last_error(const qtp_ctx *ctx);
int qtp_get_derived(const qtp_ctx *ctx, qtp_derived *out);
/* The HIP stream (hipStream_t) every call of this context runs on. */
int qtp_get_stream(const qtp_ctx *ctx, void **stream);

/* ---- reference-shaped calls (host arrays, reference layout) ----------- */

/* forces.cpp:468-630 get_forces_nospline(q, F, charge, point, list, bead):
 * zeroes and fills the bead slice of F, returns the bead energy in *energy. */
int qtp_get_forces_nospline(qtp_ctx *ctx, const double *q, double *F, const double *charge,
                            const int *point, const int *list, int bead, double *energy);

/* Pressure tensor support (no counterpart in the reference, which computes
 * no virial): get_forces_nospline plus the virial of the bead,
 *   virial[6] = W_xx, W_yy, W_zz, W_xy, W_xz, W_yz,   W = -dV/d(strain)
 * at fixed Ewald parameters (kappa, r_c, k-set), i.e. the minimum-image pair
 * sum over sites of r_ij (x) f_ij plus the intramolecular and reciprocal
 * terms.  Pressure: (N_mol k T + tr W / 3) / V for a classical system. */
int qtp_get_forces_virial(qtp_ctx *ctx, const double *q, double *F, const double *charge,
                          const int *point, const int *list, int bead, double *energy,
                          double virial[6]);

/* forces.cpp:698-789 force_eval_sc_nospline: Suzuki-Chin forces F_sc and
 * physical forces fphys for all beads; *energy = SC conserved potential,
 * *sc_pot = (2/P) sum_even V.  fphys may be NULL (as the reference default). */
int qtp_force_eval_sc_nospline(qtp_ctx *ctx, const double *q, double *F_sc,
                               const double *charge, const double *mass, const int *point,
                               const int *list, double *sc_pot, double *fphys, double *energy);

/* neighbour.cpp:12-62 neighbour_list(q, point, list, disp, bead): O-O half
 * list within r_skin, zeroes the bead's disp.  Returns QTP_ERR_LIST_OVERFLOW
 * (and writes no entry past list_size) where the reference would overflow. */
int qtp_neighbour_list(qtp_ctx *ctx, const double *q, int *point, int *list, double *disp,
                       int bead);

/* fourier.cpp:12-69 ftransform(a, length, is_forwards): in-place orthonormal
 * bead <-> normal-mode transform of a[n_beads*length]. */
int qtp_ftransform(qtp_ctx *ctx, double *a, int length, int is_forwards);

/* thermostat.cpp:64-133 pile_g(p_nm, mass, rng) with the random numbers
 * drawn by the caller, in the reference's order:
 *   dev[0] = R, dev[1] = R_add, dev[2] = gamma((dof-2)/2) draw, dev[3..] = xi
 *   for bead 1..P-1, dof 0..n_dof-1   (3 + (P-1)*n_dof values).
 * *heat receives the return value of pile_g (KE before - KE after). */
int qtp_pile_g(qtp_ctx *ctx, double *p_nm, const double *mass, const double *dev,
               double *heat);

/* update.cpp:271-401 update_sc_nospline: one B-A-O-A-B Suzuki-Chin step.
 * dev as qtp_pile_g (ignored unless is_nvt).  *heat -= pile_g's return as in
 * update.cpp:325; *energy = SC conserved potential. */
int qtp_update_sc_nospline(qtp_ctx *ctx, double *q, double *p, const double *mass, double *F,
                           double *fphys, const double *charge, int *point, int *list,
                           double *disp, const double *dev, int is_nvt, double *heat,
                           double *sc_pot, double *energy);

/* update.cpp:78-137 update_trajectory_class_nospline: one velocity-Verlet step. */
int qtp_update_trajectory_class_nospline(qtp_ctx *ctx, double *q, double *p, const double *mass,
                                         double *F, const double *charge, int *point, int *list,
                                         double *disp, double *energy);

/* path_integrals.cpp:29-48 get_kcv(q, fphys) */
int qtp_get_kcv(qtp_ctx *ctx, const double *q, const double *fphys, double *kcv);
/* get_kcv without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same estimator
 * as qtp_get_kcv (path_integrals.cpp:29-48) on DEVICE arrays in the reference
 * layout, q[b*n_dof + j] and fphys[b*n_dof + j] (n_dof = 3*n_atoms); the
 * result is written to the device double *kcv.  Asynchronous on `stream`
 * (a hipStream_t; NULL = the default stream).  No context, no
 * hipMalloc/hipFree: the block partials go to a fixed device
 * array of 8192 doubles (64 KiB per GPU, held until the process ends), so
 * n_atoms <= 699,050 (n_dof <= 2,097,152); larger inputs are rejected with
 * QTP_ERR_ARG.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (different streams, concurrent threads) race.  Errors:
 * qtp_last_error(NULL). */
int qtp_get_kcv_noalloc(int n_beads, int n_atoms, double beta, const double *q,
                        const double *fphys, double *kcv, void *stream);
/* path_integrals.cpp:52-81 get_conserved(q, p, mass, energy) (q and p are
 * not modified here; the reference round-trips them through ftransform). */
int qtp_get_conserved(qtp_ctx *ctx, const double *q, const double *p, const double *mass,
                      double energy, double *cons);

/* ---- fast-CMD splined pair potential (forces.cpp:634-691, 791-882;
 * update.cpp:16-75, 139-269) ------------------------------------------------
 * qtp_set_splines loads the reference's three spline files' contents (values
 * of the O-O, O-H and H-H pair potentials at start + j*step, j = 0..n-1;
 * simulation.cpp:36-116 reads them with start 0, step 0.01 bohr) and fits
 * them as Boost.Math's cubic_b_spline does.  Every atom pair of a type,
 * intramolecular pairs included, within start + step*(n-1) adds the spline
 * (minimum image).  After it the resident mode uses the splines until
 * qtp_clear_splines; both require qtp_set_state again.  The *_wspline calls
 * take the same arguments as their *_nospline counterparts (the reference
 * passes the spline objects and squared ranges; here they are the loaded
 * tables) and fail with QTP_ERR_STATE when no tables are loaded. */
int qtp_set_splines(qtp_ctx *ctx, const double *oo, int n_oo, const double *oh, int n_oh,
                    const double *hh, int n_hh, double start, double step);
int qtp_clear_splines(qtp_ctx *ctx);
/* forces.cpp:634-691 get_forces_wspline */
int qtp_get_forces_wspline(qtp_ctx *ctx, const double *q, double *F, const double *charge,
                           const int *point, const int *list, int bead, double *energy);
/* forces.cpp:791-882 force_eval_sc_wspline */
int qtp_force_eval_sc_wspline(qtp_ctx *ctx, const double *q, double *F_sc,
                              const double *charge, const double *mass, const int *point,
                              const int *list, double *sc_pot, double *fphys, double *energy);
/* update.cpp:139-269 update_sc_wspline */
int qtp_update_sc_wspline(qtp_ctx *ctx, double *q, double *p, const double *mass, double *F,
                          double *fphys, const double *charge, int *point, int *list,
                          double *disp, const double *dev, int is_nvt, double *heat,
                          double *sc_pot, double *energy);
/* update.cpp:16-75 update_trajectory_class_wspline */
int qtp_update_trajectory_class_wspline(qtp_ctx *ctx, double *q, double *p, const double *mass,
                                        double *F, const double *charge, int *point, int *list,
                                        double *disp, double *energy);

/* Host-side draws in the reference's order from a std::mt19937 owned by the
 * library (for callers without a C++ rng; the C++ adapter uses the caller's). */
typedef struct qtp_host_rng qtp_host_rng;
int qtp_host_rng_create(uint32_t seed, qtp_host_rng **out);
void qtp_host_rng_destroy(qtp_host_rng *r);
int qtp_host_rng_draw_pile(qtp_host_rng *r, int n_atoms, int n_beads, double *dev);

/* ---- starts and position files (host only) ------------------------------ */

/* initialise.cpp:20-98 with the box of water.cpp:90-97, 117: 4*nx*ny*nz
 * molecules on an fcc lattice at density rho_inp (g cm^-3), every bead a
 * copy of the first.  pos: n_beads * 36*nx*ny*nz doubles (reference layout);
 * boxl (may be NULL): the box edges in bohr. */
int qtp_initialise_fcc(int nx, int ny, int nz, double rho_inp, int n_beads, double *pos,
                       double *boxl);
/* initialise.cpp:189-541 with water.cpp:101-109, 117: 8*nx*ny*nz molecules
 * of proton-disordered hexagonal ice on the orthorhombic cell, satisfying the
 * ice rules with zero net dipole, drawn from rng in the reference's order
 * (initialise_orth(rng, outfile)).  max_outer > 0 bounds the number of fresh
 * hydrogen assignments (the reference loops until one has zero dipole;
 * QTP_ERR_STATE when the bound is hit); n_outer (may be NULL) returns how
 * many were made. */
int qtp_initialise_orth(qtp_host_rng *rng, int nx, int ny, int nz, double rho_inp, int n_beads,
                        long max_outer, double *pos, double *boxl, long *n_outer);
/* simulation.cpp:826-871: Maxwell-Boltzmann momenta of every bead at the
 * bead temperature n_beads*T_K (an O and an H normal_distribution drawn from
 * rng in the reference's order), minus the centre-of-mass momentum. */
int qtp_host_rng_draw_momenta(qtp_host_rng *rng, int n_mols, int n_beads, double T_K,
                              double *p);
/* initialise.cpp:107-135 file format (one atom per line, "x y z" in bohr).
 * A file of n_atoms lines is copied to every bead, as the reference does; a
 * file of n_beads*n_atoms lines (water.cpp:226-238, final_positions.txt)
 * keeps every bead, which the reference's reader does not.  beads_read (may
 * be NULL): 1 or n_beads. */
int qtp_read_positions(const char *path, int n_mols, int n_beads, double *pos, int *beads_read);
/* Every bead, one atom per line, 17 significant digits (read back exactly). */
int qtp_write_positions(const char *path, int n_mols, int n_beads, const double *pos);

/* ---- resident mode ------------------------------------------------------ */

/* Upload q and p (reference layout), charges and masses (NULL: qTIP4P/F
 * values of simulation.cpp:811-822), build every bead's list and the initial
 * forces (simulation.cpp:839-874). */
int qtp_set_state(qtp_ctx *ctx, const double *q, const double *p, const double *charge,
                  const double *mass);
/* Download state in reference layout; any pointer may be NULL. */
int qtp_get_state(qtp_ctx *ctx, double *q, double *p, double *F, double *fphys, double *disp);
/* Download the half lists in reference layout. */
int qtp_get_lists(qtp_ctx *ctx, int *point, int *list);

/* Virial of every bead at the resident positions (as qtp_get_forces_virial):
 * virial[b*6 + c]; energy[b] = V of bead b if energy is not NULL.  No
 * counterpart in the reference.  Leaves the resident state untouched. */
int qtp_get_virial(qtp_ctx *ctx, double *virial, double *energy);

/* Lossless checkpoint of the resident state (the reference restarts from
 * positions only, collapsing the beads: initialise.cpp:107-135).  Saves q,
 * p, displacements, half lists, rebuild counts, accumulated heat, the
 * thermostat seed and step counter; qtp_load_checkpoint (in place of
 * qtp_set_state; charges/masses as there) restores them into a context of
 * the same system and recomputes the forces, so the run continues exactly
 * (bit for bit) as it would have without the stop.  The file also records
 * T_K, dtfs, eps_ewald, epsilon_sc, the effective Ewald kappa, r_c, k_cut
 * and k_max, and the LJ cutoff; a load into a context where any of them (or the molecule count,
 * bead count, list size or box) differs fails with QTP_ERR_ARG. */
int qtp_save_checkpoint(qtp_ctx *ctx, const char *path);
int qtp_load_checkpoint(qtp_ctx *ctx, const char *path, const double *charge,
                        const double *mass);

/* Thermostat random numbers for resident NVT steps: Philox4x32-10 streams
 * keyed by seed and the device step counter, Box-Muller normals; the centroid
 * chi-square is a sum of squared normals.  Default seed 0. */
int qtp_set_seed(qtp_ctx *ctx, uint64_t seed);

/* Advance n_steps resident steps (SC for n_beads >= 2, velocity Verlet for
 * n_beads == 1).  is_nvt applies PILE-G (SC only).  Asynchronous: returns once
 * the work is queued; qtp_sync waits. */
int qtp_run(qtp_ctx *ctx, int n_steps, int is_nvt);
int qtp_sync(qtp_ctx *ctx);

/* Batched replicas (no counterpart in the reference, which runs one system
 * per process): n contexts, each with its own size, bead count, temperature
 * and resident state (qtp_set_state done), advanced together.  Each replica
 * keeps its own captured step graph and streams; qtp_batch_run queues every
 * replica's step, replica after replica, step after step, so the GPU
 * overlaps replicas.  is_nvt[i] (NULL: all NVE) selects PILE-G per replica.
 * While a batch runs, call no other function on its contexts before
 * qtp_batch_sync; the batch does not own the contexts (destroy it first). */
typedef struct qtp_batch qtp_batch;
int qtp_batch_create(qtp_ctx *const *ctxs, const int *is_nvt, int n, qtp_batch **out);
int qtp_batch_run(qtp_batch *b, int n_steps);
int qtp_batch_sync(qtp_batch *b);
const char *qtp_batch_last_error(const qtp_batch *b);
void qtp_batch_destroy(qtp_batch *b);

/* Scalars of the last resident step: [0] energy (SC conserved potential or
 * classical V), [1] sc_pot, [2] accumulated heat (update.cpp:325 sign),
 * [3] list rebuilds so far, [4] steps so far, [5] status flags. */
int qtp_get_scalars(qtp_ctx *ctx, double out[6]);

/* Observables of the resident state, computed on the device (no state
 * download): out[0] = K_cv (path_integrals.cpp:29-48), out[1] = ring-polymer
 * conserved quantity (get_conserved, path_integrals.cpp:52-81) of the SC
 * potential, out[2] = SC potential, out[3] = sc_pot.  For n_beads == 1:
 * kinetic energy, V + KE, V, 0.  The spring term is summed in bead space,
 * (1/2) m w_P^2 sum_b (q_{b+1} - q_b)^2, which the orthonormal transform
 * makes equal to the reference's normal-mode sum (the reference's truncated
 * pi in w_k moves it by ~1e-13 relative). */
int qtp_get_observables(qtp_ctx *ctx, double out[4]);
/* Advance n_steps resident steps recording the observables of qtp_get_observables
 * on the device after every step whose counter is a multiple of `every`;
 * trace[4 * i + c] receives sample i (at most n_steps / every + 1 samples,
 * *n_samples set).  One host synchronisation at the end (the reference
 * driver's per-sample download, simulation.cpp:951-960, goes away). */
int qtp_run_observed(qtp_ctx *ctx, int n_steps, int is_nvt, int every, double *trace,
                     int *n_samples);

/* ---- launch configuration and autotuner ------------------------------------
 * No counterpart in the reference (one CPU code path).  cfg[9] = {lanes per
 * molecule of the LJ/intramolecular kernel (16 or 64), k-space on the matrix
 * cores (1; only for k_max <= 23) or the vector kernels (0), real-space
 * coefficients in scalar registers (1), as literals (0) or by grid size (-1),
 * 16-wave k-space force blocks (1), 8-wave (0) or by block count (-1),
 * structure-factor blocks per 64-site chunk (1-4; 0 by launch size),
 * propagation as one LDS-staged kernel per step where the thermostat allows
 * it: for systems of 256 or more 64-dof tiles (-1), always (1), never (0;
 * the results are the same bit for bit), waves per real-space tile block (2
 * or 4), eighths of the real-space tiles launched on the second stream after
 * the LJ/intramolecular kernel (0-4), site chunks per structure-factor block
 * of the matrix-core path (1, 2, 4, 8, 16; 0 by launch size)}.
 * Every configuration gives the same forces up to summation order. */
int qtp_get_tuning(const qtp_ctx *ctx, int cfg[9]);
/* Sets cfg (QTP_ERR_ARG for values the context cannot run); with a state set,
 * its forces are recomputed under the new configuration. */
int qtp_set_tuning(qtp_ctx *ctx, const int cfg[9]);
/* Times one force evaluation of every bead of the resident state (the force
 * work of one step captured as a graph, replayed between HIP events, median
 * of 5 runs of replays) under each
 * candidate configuration, keeps the fastest and recomputes the state's
 * forces with it.  The choice is cached per system (molecules, beads, box,
 * Ewald and LJ parameters) for the process, so contexts of one shape are
 * timed once; flags & 1 times again.  report[11] (may be NULL): best ms,
 * default ms, the chosen cfg[0..4], candidates timed (0: from the cache),
 * the chosen cfg[6], cfg[7] and cfg[8]. */
int qtp_autotune(qtp_ctx *ctx, int flags, double report[11]);

/* ---- Verlet-list build ------------------------------------------------------
 * No counterpart in the reference, whose neighbour.cpp:12-62 tests every pair
 * and whose update.cpp:356-385 rebuilds on the displacement test only. */
#define QTP_LIST_AUTO 0      /* cells above 1024 molecules when the box holds them */
#define QTP_LIST_ALLPAIRS 1  /* every pair, as neighbour.cpp */
#define QTP_LIST_CELLS 2     /* cell-binned candidates; QTP_ERR_ARG if the box is too narrow */
/* method: how lists are built (the lists are identical either way: the same
 * FP64 pair test, rows in ascending order).  Cells are at least r_skin / S
 * wide, S = 2 when every edge holds 5 of them, else S = 1 with 3; up to
 * 65,536 molecules and 12,288 cells.  rebuild_every > 0: resident steps also
 * rebuild every bead's lists on each step whose count is a multiple of it
 * (besides the displacement test); 0 = displacement test only (default). */
int qtp_set_lists(qtp_ctx *ctx, int method, int rebuild_every);
/* info[5] = {cells in use (0/1), cells per edge x, y, z, S}; the grid is
 * reported when the box holds one, whether or not it is in use. */
int qtp_get_list_info(const qtp_ctx *ctx, int info[5]);

/* ---- centroid dynamics, thermostats, centroid mean forces, g(r) --------------
 * None of these has a counterpart in the reference, whose PILE-G is the tau -> 0
 * limit (thermostat.cpp:64-133) and which has no CMD propagator or g(r)
 * (its README leaves them to other programs). */

#define QTP_THERMO_PILE_G 0  /* centroid: global Bussi rescale, time constant tau */
#define QTP_THERMO_PILE_L 1  /* centroid: Langevin, friction 1/tau */
/* Centroid thermostat of NVT resident steps (the non-centroid modes always
 * take PILE's critically damped Langevin, thermostat.cpp:114-129).
 * PILE_G with tau_fs = 0 is pile_g exactly (the default); tau_fs > 0 the
 * canonical-sampling rescale of Bussi, Donadio and Parrinello (J. Chem.
 * Phys. 126, 014101 (2007)).  PILE_L (tau_fs > 0): Langevin on every centroid
 * degree of freedom (Ceriotti et al., J. Chem. Phys. 133, 124104 (2010)).
 * Classical contexts (n_beads == 1) take PILE_L only: a Langevin step on
 * every atom at the start of each NVT step.  The drop-in calls
 * (qtp_pile_g, qtp_update_sc_nospline) always use pile_g. */
int qtp_set_thermostat(qtp_ctx *ctx, int kind, double tau_fs);

#define QTP_MODE_PIMD 0            /* plain PIMD (default) */
#define QTP_MODE_PACMD 1           /* partially adiabatic centroid MD */
#define QTP_MODE_CENTROID_FIXED 2  /* centroid held fixed (mean-force sampling) */
/* Centroid dynamics of the resident SC steps.  PACMD gives every
 * non-centroid normal mode the frequency omega_ad (a.u.; <= 0:
 * P^(P/(P-1)) / (beta hbar), Hone, Rossky and Voth, J. Chem. Phys. 124,
 * 154103 (2006)) through the mass m (w_k / omega_ad)^2, leaves the physical
 * spring and forces as they are, and damps those modes critically; the
 * centroid keeps the physical mass.  CENTROID_FIXED zeroes the centroid
 * momentum after each kick and skips the centroid thermostat, so the beads
 * sample the ring-polymer distribution at a fixed centroid. */
int qtp_set_pimd_mode(qtp_ctx *ctx, int mode, double omega_ad);
/* on != 0: zero the accumulator and, from the next resident step on, add
 * (1/P) sum_b F_sc,b (the force on the centroid) after every step; on == 0
 * stops.  qtp_get_centroid_force returns the mean over the steps since
 * (reference layout, n_dof), optionally the current centroid positions
 * (n_dof), and the number of steps. */
int qtp_centroid_force(qtp_ctx *ctx, int on);
int qtp_get_centroid_force(qtp_ctx *ctx, double *mean_force, double *centroid, int *n_samples);
/* Pair-distance histograms of the resident (centroid, for n_beads > 1)
 * positions: O-O over molecule pairs, O-H over every (O, H) pair including
 * the molecule's own, bins of r_max / n_bins (bohr), minimum image.
 * qtp_rdf_sample queues one sample (asynchronous); qtp_rdf_get returns the
 * accumulated counts and the number of samples. */
int qtp_rdf_reset(qtp_ctx *ctx, double r_max, int n_bins);
int qtp_rdf_sample(qtp_ctx *ctx);
int qtp_rdf_get(qtp_ctx *ctx, double *oo, double *oh, int *n_samples);

/* ---- learned centroid-force correction ------------------------------------------
 * No counterpart in the reference.  It takes the place of the reference's
 * centroid hook, the additive spline pair correction of its fast-CMD path
 * (forces.cpp:452-466, 634-691): V = V_qTIP4P/F + sum_i E_i, E_i the output
 * of an element's network (O, H: 40 -> 128 -> 128 -> 1, tanh) on atom i's
 * descriptor of 24 radial Gaussians (0.8-6 A) and 16 angular terms
 * (within 4 A), each neighbour weighted by an element-pair weight.  The
 * networks run on BF16 matrix cores with FP32 accumulation; forces come back
 * in FP64 arrays.  Parameters: theta[2 * 21889] FP32, per element W1[128][40],
 * b1[128], W2[128][128], b2[128], w3[128], b3 (O first); mean[2][40] and
 * std[2][40] normalise the descriptors; E_i carries the factor e_scale. */
typedef struct qtp_nn qtp_nn;
/* The descriptor constants (r_rad, r_ang, mu0, dmu, eta_r, zeta[16],
 * lambda[16], eta_a[16], w[2][2]; Angstrom): writes min(n, count) values,
 * returns the count (57).  No GPU needed. */
int qtp_nn_descriptor_params(double *out, int n);
/* A model for systems of n_mols waters in the box boxl (bohr; every edge
 * above 12 A), with work space for batches of up to max_frames frames. */
int qtp_nn_create(int n_mols, const double boxl[3], int max_frames, qtp_nn **out);
void qtp_nn_destroy(qtp_nn *nn);
const char *qtp_nn_last_error(const qtp_nn *nn);
int qtp_nn_set_model(qtp_nn *nn, const float *theta, const float *mean, const float *stdv,
                     double e_scale);
int qtp_nn_get_model(qtp_nn *nn, float *theta);
/* Energies (n_frames), forces (n_frames * n_dof, Hartree/bohr) and raw
 * descriptors (n_frames * n_atoms * 40) of frames in the reference layout;
 * any output may be NULL. */
int qtp_nn_eval(qtp_nn *nn, int n_frames, const double *pos, double *energy, double *forces,
                float *descriptors);
/* Training data for force matching: positions and target forces (both
 * n_frames * n_dof, reference layout), kept on the device.  The loss is
 * the mean over atoms and components of (F_model - target)^2. */
int qtp_nn_set_data(qtp_nn *nn, int n_frames, const double *pos, const double *target);
/* Loss and its gradient in theta (2 * 21889) on frames [frame0, frame0 +
 * n_frames) of the data (n_frames <= max_frames), no update. */
int qtp_nn_loss_grad(qtp_nn *nn, int frame0, int n_frames, double *loss, float *grad);
/* n_steps Adam steps (beta 0.9, 0.999) with learning rate lr on batches of
 * batch_frames consecutive data frames, batches in a seeded random order per
 * pass; loss_trace (may be NULL) receives each batch's loss before its step. */
int qtp_nn_train(qtp_nn *nn, int n_steps, int batch_frames, double lr, uint64_t seed,
                 double *loss_trace);
/* Attach (nn) or detach (NULL) the correction to a classical context's
 * resident steps: its forces and energy are added to qTIP4P/F's (the model
 * must be made for the same molecule count and box).  qtp_set_state must
 * follow. */
int qtp_set_nn(qtp_ctx *ctx, qtp_nn *nn);

/* Per-kernel timing of the last qtp_run (HIP events around each launch of the
 * named kernel class on the context's stream); for measurement only. */
int qtp_enable_timing(qtp_ctx *ctx, int on);
int qtp_get_timing(qtp_ctx *ctx, const char *kernel, double *ms_total, int *launches);

/* The Philox4x32-10 block function the device thermostat uses (host build of
 * the same code), for known-answer checks.  No GPU needed. */
int qtp_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
