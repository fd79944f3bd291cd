// qtp_setup.cpp -- host-side starts and position files: the lattices of
// initialise.cpp, the initial momenta of simulation.cpp and the position file
// format, with the reference's arithmetic and random-number order so that a
// seeded start is the reference's start.  No device code.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../../include/qtip4pf.h"
#include "qtp_host.hpp"

using namespace qtp;

namespace {

// qtip4pfparams.h:49, 64-68, 95, in the reference's order of evaluation
struct LatticeConsts {
    double mh2o, one_third, two_thirds, one_sixth, five_sixths, sqrt3, zo;
    LatticeConsts()
    {
        const double mhgmol = 1.0080, mogmol = 15.9994;
        mh2o = (mhgmol * 2.0 + mogmol) / RC.Na;
        one_third = 1.0 / 3.0;
        two_thirds = 2.0 * one_third;
        one_sixth = 0.5 * one_third;
        five_sixths = 5.0 / 6.0;
        sqrt3 = std::sqrt(3.0);
        zo = 1.0 / 16.0;
    }
};
const LatticeConsts LC;

// Minimum-image wrap of one component (initialise.cpp:155-160, :352-357).
inline double wrap(double r, double L)
{
    if (r <= -0.5 * L)
        r += L;
    else if (r > 0.5 * L)
        r -= L;
    return r;
}

bool dims_ok(int nx, int ny, int nz, double rho, int n_beads)
{
    return nx > 0 && ny > 0 && nz > 0 && n_beads > 0 && rho > 0.0 && std::isfinite(rho) &&
           (long)nx * ny * nz <= (1L << 20);
}

void copy_bead0(double *pos, int n_beads, long n_dof)
{
    for (int b = 1; b < n_beads; ++b)
        for (long i = 0; i < n_dof; ++i) pos[b * n_dof + i] = pos[i];
}

// One lattice coordinate ((cell + s1) + s2) * edge: the two shifts are added in
// the order the reference's expressions add them (e.g. (k + 0.5 - zo) * c_cell).
struct Site {
    double s1[3], s2[3];
};

inline double coord(int cell, const Site &st, int d, double edge)
{
    return ((cell + st.s1[d]) + st.s2[d]) * edge;
}

// A bond of the orthorhombic ice lattice: site a of a cell to site b of the
// cell displaced by (di, dj, dk), periodic (initialise.cpp:255-332).
struct Bond {
    int a, b, di, dj, dk;
};
const Bond ORTH_BONDS[16] = {
    {0, 1, 0, 0, 0},  {0, 3, 0, 0, -1}, {0, 3, 1, 0, -1}, {0, 7, 0, -1, -1},
    {1, 2, 0, 0, 0},  {1, 2, 1, 0, 0},  {1, 6, 0, -1, 0}, {2, 3, 0, 0, 0},
    {2, 5, 0, 0, 0},  {3, 4, 0, 0, 1},  {4, 5, 0, 0, 0},  {4, 7, 0, 0, -1},
    {4, 7, -1, 0, -1}, {5, 6, 0, 0, 0}, {5, 6, -1, 0, 0}, {6, 7, 0, 0, 0},
};

}  // namespace

extern "C" {

int qtp_initialise_fcc(int nx, int ny, int nz, double rho_inp, int n_beads, double *pos,
                       double *boxl)
{
    if (!pos || !dims_ok(nx, ny, nz, rho_inp, n_beads)) return QTP_ERR_ARG;
    // water.cpp:93-97, :117
    const double a = 1e8 * std::cbrt(4 * LC.mh2o / rho_inp) * RC.angs2bohr;
    if (boxl) {
        boxl[0] = nx * a;
        boxl[1] = ny * a;
        boxl[2] = nz * a;
    }
    // initialise.cpp:28-29
    const double xdiff = RC.roheq * std::cos(RC.theta_eq / 2.0) / std::sqrt(2.0);
    const double zdiff = RC.roheq * std::sin(RC.theta_eq / 2.0);
    // the four oxygens of a cell in cell units (initialise.cpp:39-77)
    static const double F[4][3] = {
        {0.25, 0.25, 0.25}, {0.25, 0.75, 0.75}, {0.75, 0.25, 0.75}, {0.75, 0.75, 0.25}};
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const long cell = k + (long)nz * j + (long)nz * ny * i;
                for (int s = 0; s < 4; ++s) {
                    double *m = pos + 9 * (4 * cell + s);
                    const int idx[3] = {i, j, k};
                    for (int d = 0; d < 3; ++d) m[d] = (idx[d] + F[s][d]) * a;
                    // both hydrogens at +xdiff in x and y, +-zdiff in z
                    m[3] = m[0] + xdiff;
                    m[4] = m[1] + xdiff;
                    m[5] = m[2] + zdiff;
                    m[6] = m[0] + xdiff;
                    m[7] = m[1] + xdiff;
                    m[8] = m[2] - zdiff;
                }
            }
    copy_bead0(pos, n_beads, 36L * nx * ny * nz);
    return QTP_OK;
}

int qtp_initialise_orth(qtp_host_rng *rng, int nx, int ny, int nz, double rho_inp, int n_beads,
                        long max_outer, double *pos, double *boxl, long *n_outer)
{
    if (!rng || !pos || !dims_ok(nx, ny, nz, rho_inp, n_beads)) return QTP_ERR_ARG;
    // water.cpp:104-109, :117
    const double a = 1e8 * std::cbrt(4 * LC.mh2o / (std::sqrt(2.0) * rho_inp)) * RC.angs2bohr;
    const double edge[3] = {a, LC.sqrt3 * a, 2.0 * std::sqrt(2.0 / 3.0) * a};
    const double rnn = (0.5 - 2.0 * LC.zo) * edge[2];
    const double L[3] = {nx * edge[0], ny * edge[1], nz * edge[2]};
    if (boxl)
        for (int d = 0; d < 3; ++d) boxl[d] = L[d];
    const int n_cells = nx * ny * nz, n_mols = 8 * n_cells, n_bonds = 2 * n_mols;
    const long n_dof = 9L * n_mols;

    // oxygens on two interleaved hexagonal sublattices (initialise.cpp:189-250)
    const double zo = LC.zo;
    const Site sites[8] = {
        {{0.5, LC.one_sixth, zo}, {0.0, 0.0, 0.0}},
        {{0.5, LC.one_sixth, 0.5}, {0.0, 0.0, -zo}},
        {{0.0, LC.one_third, 0.5}, {0.0, 0.0, zo}},
        {{0.0, LC.one_third, 1.0}, {0.0, 0.0, -zo}},
        {{0.0, LC.two_thirds, zo}, {0.0, 0.0, 0.0}},
        {{0.0, LC.two_thirds, 0.5}, {0.0, 0.0, -zo}},
        {{0.5, LC.five_sixths, 0.5}, {0.0, 0.0, zo}},
        {{0.5, LC.five_sixths, 1.0}, {0.0, 0.0, -zo}},
    };
    for (long t = 0; t < n_dof; ++t) pos[t] = 0.0;
    std::vector<int> bond_a(n_bonds), bond_b(n_bonds);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int cell = k + nz * j + nz * ny * i;
                const int idx[3] = {i, j, k};
                for (int s = 0; s < 8; ++s)
                    for (int d = 0; d < 3; ++d) pos[9 * (8 * cell + s) + d] = coord(idx[d], sites[s], d, edge[d]);
                for (int e = 0; e < 16; ++e) {
                    const Bond &B = ORTH_BONDS[e];
                    const int i2 = (i + B.di + nx) % nx, j2 = (j + B.dj + ny) % ny,
                              k2 = (k + B.dk + nz) % nz;
                    bond_a[16 * cell + e] = 8 * cell + B.a;
                    bond_b[16 * cell + e] = 8 * (k2 + nz * j2 + nz * ny * i2) + B.b;
                }
            }

    // unit O(a)->O(b) vector scaled by the nearest-neighbour distance
    auto unit_oo = [&](int e, double u[3]) {
        for (int d = 0; d < 3; ++d)
            u[d] = wrap(pos[9 * bond_a[e] + d] - pos[9 * bond_b[e] + d], L[d]) / rnn;
    };
    // hydrogen of bond e on oxygen a (pulled back along the bond) or b
    auto place = [&](int e, bool on_a, double *h) {
        double u[3];
        unit_oo(e, u);
        for (int d = 0; d < 3; ++d)
            h[d] = on_a ? pos[9 * bond_a[e] + d] - RC.roheq * u[d]
                        : pos[9 * bond_b[e] + d] + RC.roheq * u[d];
    };

    std::uniform_int_distribution<int> coin(0, 1);
    std::uniform_int_distribution<int> pick(0, n_bonds - 1);
    std::vector<int> cn(n_mols), owner(n_bonds), slot(n_mols);
    std::vector<double> hyd(3 * (size_t)n_bonds);
    long outer = 0;
    for (;;) {
        if (max_outer > 0 && outer >= max_outer) return QTP_ERR_STATE;
        ++outer;
        // one hydrogen per O-O bond, on either end (initialise.cpp:338-376)
        std::fill(cn.begin(), cn.end(), 0);
        for (int e = 0; e < n_bonds; ++e) {
            const bool on_a = coin(rng->gen) != 0;
            place(e, on_a, &hyd[3 * e]);
            owner[e] = on_a ? bond_a[e] : bond_b[e];
            cn[owner[e]] += 1;
        }
        // move hydrogens across bonds until every oxygen holds two: a move
        // that lowers |c_a - c_b| is taken, one that keeps it is taken when a
        // second coin shows 0 (initialise.cpp:378-460)
        for (;;) {
            bool done = true;
            for (int m = 0; m < n_mols; ++m) done = done && cn[m] == 2;
            if (done) break;
            const int e = pick(rng->gen);
            const int oa = bond_a[e], ob = bond_b[e];
            const int old_diff = std::abs(cn[oa] - cn[ob]);
            const bool from_a = owner[e] == oa;
            const int new_diff = std::abs(cn[oa] - cn[ob] + (from_a ? -2 : 2));
            if (new_diff < old_diff || (new_diff == old_diff && coin(rng->gen) == 0)) {
                place(e, !from_a, &hyd[3 * e]);
                owner[e] = from_a ? ob : oa;
                cn[owner[e]] += 1;
                cn[from_a ? oa : ob] -= 1;
            }
        }
        // hydrogens into the slots after their oxygen, in bond order
        // (initialise.cpp:462-479)
        std::fill(slot.begin(), slot.end(), 0);
        for (int e = 0; e < n_bonds; ++e) {
            const int o = owner[e];
            slot[o] += 1;
            for (int d = 0; d < 3; ++d) pos[9 * o + 3 * slot[o] + d] = hyd[3 * e + d];
        }
        // sum of the unit H-O-H bisectors (initialise.cpp:137-185) must vanish
        double dip[3] = {0.0, 0.0, 0.0};
        for (int m = 0; m < n_mols; ++m) {
            const double *M = pos + 9 * m;
            double bis[3];
            for (int d = 0; d < 3; ++d) bis[d] = wrap(M[3 + d] - M[d], L[d]) + wrap(M[6 + d] - M[d], L[d]);
            const double rb = std::sqrt(bis[0] * bis[0] + bis[1] * bis[1] + bis[2] * bis[2]);
            for (int d = 0; d < 3; ++d) dip[d] += bis[d] / rb;
        }
        if (std::fabs(dip[0]) < 1e-10 && std::fabs(dip[1]) < 1e-10 && std::fabs(dip[2]) < 1e-10) break;
    }
    if (n_outer) *n_outer = outer;
    copy_bead0(pos, n_beads, n_dof);
    return QTP_OK;
}

int qtp_host_rng_draw_momenta(qtp_host_rng *rng, int n_mols, int n_beads, double T_K, double *p)
{
    if (!rng || !p || n_mols < 1 || n_beads < 1 || !(T_K > 0.0)) return QTP_ERR_ARG;
    // water.cpp:77-80, 138-139
    const double beta = 1.0 / (RC.K2au * T_K);
    const double beta_n = beta / n_beads;
    std::normal_distribution<double> ox(0.0, std::sqrt(RC.mo / beta_n));
    std::normal_distribution<double> hy(0.0, std::sqrt(RC.mh / beta_n));
    const long n_dof = 9L * n_mols;
    double pc[3] = {0.0, 0.0, 0.0};
    // simulation.cpp:848-857: per bead, molecule and axis: O, H1, H2
    for (int b = 0; b < n_beads; ++b)
        for (int m = 0; m < n_mols; ++m)
            for (int d = 0; d < 3; ++d) {
                double *M = p + b * n_dof + 9L * m;
                M[d] = ox(rng->gen);
                M[d + 3] = hy(rng->gen);
                M[d + 6] = hy(rng->gen);
                pc[d] += M[d] + M[d + 3] + M[d + 6];
            }
    // simulation.cpp:859-870: remove the centre-of-mass momentum of the
    // whole ring-polymer system
    for (int d = 0; d < 3; ++d) pc[d] /= 3 * n_mols * n_beads;
    for (int b = 0; b < n_beads; ++b)
        for (long a = 0; a < 3L * n_mols; ++a)
            for (int d = 0; d < 3; ++d) p[b * n_dof + 3 * a + d] -= pc[d];
    return QTP_OK;
}

int qtp_read_positions(const char *path, int n_mols, int n_beads, double *pos, int *beads_read)
{
    if (!path || !pos || n_mols < 1 || n_beads < 1) return QTP_ERR_ARG;
    FILE *f = std::fopen(path, "r");
    if (!f) return QTP_ERR_ARG;
    const long n_atoms = 3L * n_mols, cap = n_atoms * n_beads;
    long count = 0;
    char line[4096];
    bool bad = false;
    // initialise.cpp:107-126: one atom per non-empty line, "x y z" in bohr
    while (std::fgets(line, sizeof line, f)) {
        char *s = line;
        while (*s == ' ' || *s == '\t' || *s == '\r' || *s == '\n') ++s;
        if (!*s) continue;
        if (count >= cap) {
            bad = true;
            break;
        }
        for (int d = 0; d < 3; ++d) {
            char *e = nullptr;
            const double v = std::strtod(s, &e);
            if (e == s) {
                bad = true;
                break;
            }
            pos[3 * count + d] = v;
            s = e;
        }
        if (bad) break;
        ++count;
    }
    std::fclose(f);
    // a file of one bead is copied to every bead, as the reference does
    // (initialise.cpp:128-132); a file of every bead keeps each bead
    int nb = 0;
    if (!bad && count == n_atoms)
        nb = 1;
    else if (!bad && count == cap)
        nb = n_beads;
    if (nb == 0) return QTP_ERR_ARG;
    if (nb == 1) copy_bead0(pos, n_beads, 3 * n_atoms);
    if (beads_read) *beads_read = nb;
    return QTP_OK;
}

int qtp_write_positions(const char *path, int n_mols, int n_beads, const double *pos)
{
    if (!path || !pos || n_mols < 1 || n_beads < 1) return QTP_ERR_ARG;
    FILE *f = std::fopen(path, "w");
    if (!f) return QTP_ERR_ARG;
    // water.cpp:226-238: every bead, one atom per line; 17 significant
    // digits so that reading the file back is exact
    const long n = 3L * n_mols * n_beads;
    bool ok = true;
    for (long a = 0; a < n && ok; ++a)
        ok = std::fprintf(f, "%.17g %.17g %.17g\n", pos[3 * a], pos[3 * a + 1], pos[3 * a + 2]) > 0;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? QTP_OK : QTP_ERR_ARG;
}

}  // extern "C"
