// adapter_driver.cpp -- test driver in the role of the reference's
// run_sc_nospline / run_classical_nospline (simulation.cpp:306-460, 795-998):
// it defines the process globals (water.cpp:29-38), builds the lists, takes
// the initial forces and advances the system through the reference-named
// functions of include/qtip4pf_ref_adapter.hpp.
//
// usage: qtp_adapter_driver <in.bin> <out.bin>
//   in.bin : int32 n_mols, n_beads, n_steps, is_nvt, seed;
//            float64 L[3], T_K, dtfs, eps, epsilonsc;
//            float64 q[P*n_dof], p[P*n_dof]
//   out.bin: float64 q[P*n_dof], p[P*n_dof], F[P*n_dof],
//            energy[n_steps], heat, sc_pot
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

// water.cpp:29-38
double T_K, epsilonsc, rho_inp, dtfs, eps, a_cell_angs, a_cell, b_cell, c_cell, rnn;
int n_beads, n_cell_x, n_cell_y, n_cell_z, n_cells;
int n_mols, n_atoms, n_dof, list_size;
double T_red, beta, piovern, beta_n, omega_p, omega_psq;
double v_cell, vol, rho_ox, sd_ox, sd_h, dt, halfdt;
double acc, rho, min_xy, min_box, kwald_prefactor, max_xy, max_box;
std::array<double, 3> boxl;

#define QTIP4PF_ADAPTER_IMPLEMENTATION
#include "qtip4pf_ref_adapter.hpp"

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
        return 2;
    }
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int hdr[5];
    double par[7];
    if (std::fread(hdr, sizeof(int), 5, f) != 5 || std::fread(par, sizeof(double), 7, f) != 7) return 2;
    n_mols = hdr[0];
    n_beads = hdr[1];
    const int n_steps = hdr[2], is_nvt = hdr[3];
    const unsigned seed = (unsigned)hdr[4];
    boxl = {par[0], par[1], par[2]};
    T_K = par[3];
    dtfs = par[4];
    eps = par[5];
    epsilonsc = par[6];
    n_atoms = 3 * n_mols;
    n_dof = 3 * n_atoms;
    // derived as water.cpp:135 with rho_ox = n_mols / V (cubic box given directly)
    const double r_skin = 1.30 * 9.0 * (1.0 / 0.529177210903);
    vol = boxl[0] * boxl[1] * boxl[2];
    rho_ox = n_mols / vol;
    list_size = (int)std::round(4.0 * r_skin * r_skin * r_skin * n_mols * rho_ox);
    const size_t N = (size_t)n_beads * n_dof;
    std::vector<double> q(N), p(N), F(N), fphys(N), disp((size_t)n_beads * 3 * n_mols, 0.0);
    if (std::fread(q.data(), sizeof(double), N, f) != N || std::fread(p.data(), sizeof(double), N, f) != N)
        return 2;
    std::fclose(f);
    std::vector<int> point((size_t)n_beads * n_mols), list((size_t)n_beads * list_size);
    std::vector<double> charge(n_atoms), mass(n_dof);
    for (int i = 0; i < n_mols; ++i) {  // simulation.cpp:811-822
        charge[3 * i] = -1.1128;
        charge[3 * i + 1] = charge[3 * i + 2] = 0.5564;
        const double mo = 0.001 * 15.9994 / (6.0221409e23 * 9.10938356e-31);
        const double mh = 0.001 * 1.0080 / (6.0221409e23 * 9.10938356e-31);
        for (int d = 0; d < 3; ++d) {
            mass[9 * i + d] = mo;
            mass[9 * i + d + 3] = mass[9 * i + d + 6] = mh;
        }
    }
    for (int b = 0; b < n_beads; ++b) neighbour_list(q.data(), point.data(), list.data(), disp.data(), b);
    std::mt19937 rng(seed);
    double heat = 0.0, sc_pot = 0.0;
    std::vector<double> energy(n_steps);
    if (n_beads > 1) {
        force_eval_sc_nospline(q.data(), F.data(), charge.data(), mass.data(), point.data(),
                               list.data(), sc_pot, fphys.data());
        for (int s = 0; s < n_steps; ++s)
            energy[s] = update_sc_nospline(q.data(), p.data(), mass.data(), F.data(), fphys.data(),
                                           charge.data(), point.data(), list.data(), disp.data(),
                                           rng, is_nvt != 0, heat, sc_pot);
    } else {
        get_forces_nospline(q.data(), F.data(), charge.data(), point.data(), list.data(), 0);
        for (int s = 0; s < n_steps; ++s)
            energy[s] = update_trajectory_class_nospline(q.data(), p.data(), mass.data(), F.data(),
                                                         charge.data(), point.data(), list.data(),
                                                         disp.data());
    }
    FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    std::fwrite(q.data(), sizeof(double), N, o);
    std::fwrite(p.data(), sizeof(double), N, o);
    std::fwrite(F.data(), sizeof(double), N, o);
    std::fwrite(energy.data(), sizeof(double), n_steps, o);
    std::fwrite(&heat, sizeof(double), 1, o);
    std::fwrite(&sc_pot, sizeof(double), 1, o);
    std::fclose(o);
    return 0;
}
