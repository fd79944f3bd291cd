// ref_harness.cpp -- harness around the reference's own neighbour.cpp and
// thermostat.cpp (and path_integrals.cpp's get_kcv), compiled unmodified
// from the reference tree (see Makefile).
//
// TEST INFRASTRUCTURE ONLY.  The reference keeps its state in process globals
// defined by water.cpp:29-38; water.cpp itself needs FFTW and Boost headers
// this image lacks, so this harness (the caller, in the role of water.cpp's
// main) defines those globals and exposes the two routines over extern "C"
// for the tests.  forces.cpp, update.cpp and fourier.cpp need the missing
// Boost/FFTW headers and are not built.
#include <array>
#include <random>
#include <string>
#include <cstring>
#include <cmath>

#include "qtip4pfparams.h"
#include "neighbour.h"
#include "path_integrals.h"
#include "thermostat.h"
#include "indexing.h"

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

}  // extern "C"
