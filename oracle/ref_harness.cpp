// ref_harness.cpp -- harness around the reference's own neighbour.cpp,
// thermostat.cpp, initialise.cpp (and path_integrals.cpp's get_kcv), compiled
// unmodified from the reference tree (see Makefile).
//
// TEST INFRASTRUCTURE ONLY.  The reference keeps its state in process globals
// defined by water.cpp:29-38; water.cpp itself needs FFTW and Boost headers
// this image lacks, so this harness (the caller, in the role of water.cpp's
// main) defines those globals and exposes the two routines over extern "C"
// for the tests.  forces.cpp, update.cpp and fourier.cpp need the missing
// Boost/FFTW headers and are not built.
#include <array>
#include <fstream>
#include <random>
#include <string>
#include <cstring>
#include <cmath>
#include <vector>

#include "qtip4pfparams.h"
#include "neighbour.h"
#include "path_integrals.h"
#include "thermostat.h"
#include "indexing.h"
#include "initialise.h"

// water.cpp:29-38
double T_K, epsilonsc, rho_inp, dtfs, eps, a_cell_angs, a_cell, b_cell, c_cell, rnn;
int n_beads, n_cell_x, n_cell_y, n_cell_z, n_cells;
int n_mols, n_atoms, n_dof, list_size;
double T_red, beta, piovern, beta_n, omega_p, omega_psq;
double v_cell, vol, rho_ox, sd_ox, sd_h, dt, halfdt;
double acc, rho, min_xy, min_box, kwald_prefactor, max_xy, max_box;
std::array<double, 3> boxl;
int n_equil, n_samp;
bool is_spline;

static std::mt19937 g_rng;

extern "C" {

// Set the globals the two compiled routines read.
void ref_set_globals(int n_mols_, int n_beads_, const double *boxl_, int list_size_,
                     double beta_, double beta_n_, double dt_)
{
    n_mols = n_mols_;
    n_atoms = 3 * n_mols;
    n_dof = 3 * n_atoms;
    n_beads = n_beads_;
    boxl = {boxl_[0], boxl_[1], boxl_[2]};
    list_size = list_size_;
    beta = beta_;
    beta_n = beta_n_;
    dt = dt_;
    halfdt = 0.5 * dt;
}

void ref_seed(unsigned int seed) { g_rng.seed(seed); }

// path_integrals.cpp:29-48 (reads n_beads, n_dof, beta)
double ref_get_kcv(const double *q, const double *f) { return get_kcv(q, f); }

// neighbour.cpp:12-62
void ref_neighbour_list(const double *q, int *point, int *list, double *disp, int bead)
{
    neighbour_list(q, point, list, disp, bead);
}

// thermostat.cpp:64-133 with the harness's mt19937
double ref_pile_g(double *p, const double *mass) { return pile_g(p, mass, g_rng); }

// thermostat.cpp:12-60 (not on the hot path; kept for completeness of the check)
double ref_bussi(double *p, const double *mass) { return bussi(p, mass, g_rng); }

// Draw the deviates pile_g would consume from the harness rng, in its order
// (thermostat.cpp:71, 88-98, 124): one normal_distribution object serves R,
// R_add and every xi; the gamma distribution is a separate object.
void ref_draw_pile_deviates(double *dev)
{
    std::normal_distribution<double> gauss_dev(0.0, 1.0);
    int dof = 3 * n_atoms;
    dev[0] = gauss_dev(g_rng);
    if (dof % 2 == 0) {
        std::gamma_distribution<double> gamma_dev((dof - 2) / 2, 1.0);
        dev[1] = gauss_dev(g_rng);
        dev[2] = gamma_dev(g_rng);
    } else {
        std::gamma_distribution<double> gamma_dev((dof - 1) / 2, 1.0);
        dev[1] = 0.0;
        dev[2] = gamma_dev(g_rng);
    }
    size_t n = (size_t)(n_beads - 1) * n_dof;
    for (size_t i = 0; i < n; ++i) dev[3 + i] = gauss_dev(g_rng);
}

// water.cpp:84-118: the lattice globals initialise.cpp reads, for cell type
// fcc (orth = 0) or orth (orth = 1); returns the box.
void ref_set_lattice(int orth, int nx, int ny, int nz, double rho, int n_beads_, double *box)
{
    n_cell_x = nx;
    n_cell_y = ny;
    n_cell_z = nz;
    rho_inp = rho;
    n_beads = n_beads_;
    n_cells = n_cell_x * n_cell_y * n_cell_z;
    if (!orth) {
        n_mols = 4 * n_cells;
        a_cell_angs = 1e8 * std::cbrt(4 * mh2o / rho_inp);
        a_cell = a_cell_angs * angs2bohr;
        b_cell = a_cell;
        c_cell = a_cell;
    } else {
        n_mols = 8 * n_cells;
        a_cell_angs = 1e8 * std::cbrt(4 * mh2o / (sqrt(2.0) * rho_inp));
        a_cell = a_cell_angs * angs2bohr;
        b_cell = sqrt3 * a_cell;
        c_cell = 2.0 * sqrt(2.0 / 3.0) * a_cell;
        rnn = (0.5 - 2.0 * zo) * c_cell;
    }
    n_atoms = 3 * n_mols;
    n_dof = 3 * n_atoms;
    boxl = {n_cell_x * a_cell, n_cell_y * b_cell, n_cell_z * c_cell};
    for (int d = 0; d < 3; ++d) box[d] = boxl[d];
}

// initialise.cpp:20-98 (fcc) or :334-541 (orth, with the harness rng; its log
// lines are discarded).  out: n_beads * n_dof.
void ref_initialise(int orth, double *out)
{
    double *pa;
    if (orth) {
        std::ofstream log("/dev/null");
        pa = initialise_orth(g_rng, log);
    } else {
        pa = initialise_fcc();
    }
    std::memcpy(out, pa, sizeof(double) * (size_t)n_beads * n_dof);
    delete[] pa;
}

// initialise.cpp:107-135 on a file.  out: n_beads * n_dof.
void ref_initialise_fromfile(const char *path, double *out)
{
    double *pa = initialise_fromfile(std::string(path));
    std::memcpy(out, pa, sizeof(double) * (size_t)n_beads * n_dof);
    delete[] pa;
}

// simulation.cpp:826-871 (inside a routine that also needs forces.cpp, so
// restated here): bead momenta from two normal distributions, then the
// centre-of-mass momentum removed.  Reads n_mols, n_beads, beta_n.
void ref_draw_momenta(double *p)
{
    sd_ox = sqrt(mo / beta_n);
    sd_h = sqrt(mh / beta_n);
    std::normal_distribution<double> boltz_dev_ox(0.0, sd_ox);
    std::normal_distribution<double> boltz_dev_h(0.0, sd_h);
    std::array<double, 3> p_cofm = {0.0, 0.0, 0.0};
    for (int bead = 0; bead < n_beads; ++bead)
        for (int i = 0; i < n_mols; ++i)
            for (int dim = 0; dim < 3; ++dim) {
                p[bindex(bead, 9 * i + dim)] = boltz_dev_ox(g_rng);
                p[bindex(bead, 9 * i + dim + 3)] = boltz_dev_h(g_rng);
                p[bindex(bead, 9 * i + dim + 6)] = boltz_dev_h(g_rng);
                p_cofm[dim] += p[bindex(bead, 9 * i + dim)] + p[bindex(bead, 9 * i + dim + 3)] +
                               p[bindex(bead, 9 * i + dim + 6)];
            }
    for (int dim = 0; dim < 3; ++dim) p_cofm[dim] /= n_atoms * n_beads;
    for (int bead = 0; bead < n_beads; ++bead)
        for (int i = 0; i < n_atoms; ++i)
            for (int dim = 0; dim < 3; ++dim) p[bindex(bead, 3 * i + dim)] -= p_cofm[dim];
}

void ref_set_beta_n(double beta_n_) { beta_n = beta_n_; }

}  // extern "C"
