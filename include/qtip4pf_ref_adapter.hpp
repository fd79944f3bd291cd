// qtip4pf_ref_adapter.hpp -- the reference driver's hot-path symbols over the
// C ABI of qtip4pf.h (libqtip4pf.so, MI355X / gfx950).
//
// Drop-in for rayseashyn/qtip4pf_water: build water.cpp, simulation.cpp and
// initialise.cpp unchanged, drop forces.cpp, neighbour.cpp, fourier.cpp,
// update.cpp, thermostat.cpp and path_integrals.cpp from the build, and add
// ONE translation unit containing
//
//     #include "qtip4pfparams.h"          // the reference's globals
//     #define QTIP4PF_ADAPTER_IMPLEMENTATION
//     #include "qtip4pf_ref_adapter.hpp"
//
// then link with -lqtip4pf.  Each function below has the reference's exact
// signature (default arguments stay in the reference headers) and argument
// meaning; the process globals of water.cpp:29-38 configure a lazily created
// context on first use.  The spline (fast-CMD) variants are defined when
// Boost.Math's cubic_b_spline header is available (the reference needs it
// for them anyway): the tables are read back from the caller's spline
// objects at their knots (start 0, step 0.01 bohr, simulation.cpp:40-41,
// 66-67, 93-94) and loaded once per set of spline objects.
//
// Random numbers: pile_g and update_sc_nospline(is_nvt) draw from the
// caller's std::mt19937 with libstdc++'s normal_distribution and
// gamma_distribution in the reference's order (thermostat.cpp:71, 88-98,
// 124), so a run with the same seed consumes the same stream.
//
// Errors: the reference reports problems on std::cerr and continues
// (neighbour.cpp:57); the adapter does the same with the library's message.
#ifndef QTIP4PF_REF_ADAPTER_HPP
#define QTIP4PF_REF_ADAPTER_HPP

#include <array>
#include <random>

#include "qtip4pf.h"

// Reference entry points (forces.h:16,20; neighbour.h:4; fourier.h:4;
// update.h:5,7; thermostat.h:6; path_integrals.h:5-6).
double get_forces_nospline(const double position_array[], double force_array[],
                           const double charge_array[], const int point[], const int list[],
                           int bead);
double force_eval_sc_nospline(const double position_array[], double force_array[],
                              const double charge_array[], const double mass_array[],
                              const int point[], const int list[], double &sc_pot,
                              double fphys[]);
void neighbour_list(const double position_array[], int point[], int list[], double disp_array[],
                    int bead);
void ftransform(double input_array[], const int length, const bool is_forwards);
double pile_g(double momentum_array[], const double mass_array[], std::mt19937 &gen);
double update_trajectory_class_nospline(double position_array[], double momentum_array[],
                                        const double mass_array[], double force_array[],
                                        double charge_array[], int point[], int list[],
                                        double disp_array[]);
double update_sc_nospline(double position_array[], double momentum_array[],
                          const double mass_array[], double force_array[], double fphys[],
                          double charge_array[], int point[], int list[], double disp_array[],
                          std::mt19937 &rng, bool is_nvt, double &heat, double &sc_pot);
double get_kcv(const double position_array[], const double force_array[]);
double get_conserved(double position_array[], double momentum_array[], const double mass_array[],
                     double energy);

#if defined(__has_include)
#if __has_include(<boost/math/interpolators/cubic_b_spline.hpp>)
#include <boost/math/interpolators/cubic_b_spline.hpp>
#define QTIP4PF_ADAPTER_SPLINES 1
// forces.h:18, 22; update.h:4, 6
double get_forces_wspline(const double position_array[], double force_array[],
                          const double charge_array[], const int point[], const int list[],
                          const boost::math::cubic_b_spline<double> &oo_spline,
                          const boost::math::cubic_b_spline<double> &oh_spline,
                          const boost::math::cubic_b_spline<double> &hh_spline,
                          const double oo_spline_cutsq, const double oh_spline_cutsq,
                          const double hh_spline_cutsq, int bead);
double force_eval_sc_wspline(const double position_array[], double force_array[],
                             const double charge_array[], const double mass_array[],
                             const int point[], const int list[],
                             const boost::math::cubic_b_spline<double> &oo_spline,
                             const boost::math::cubic_b_spline<double> &oh_spline,
                             const boost::math::cubic_b_spline<double> &hh_spline,
                             const double oo_spline_cutsq, const double oh_spline_cutsq,
                             const double hh_spline_cutsq, double &sc_pot, double fphys[]);
double update_trajectory_class_wspline(double position_array[], double momentum_array[],
                                       const double mass_array[], double force_array[],
                                       double charge_array[], int point[], int list[],
                                       double disp_array[],
                                       const boost::math::cubic_b_spline<double> &oo_spline,
                                       const boost::math::cubic_b_spline<double> &oh_spline,
                                       const boost::math::cubic_b_spline<double> &hh_spline,
                                       const double oo_spline_cutsq, const double oh_spline_cutsq,
                                       const double hh_spline_cutsq);
double update_sc_wspline(double position_array[], double momentum_array[],
                         const double mass_array[], double force_array[], double fphys[],
                         double charge_array[], int point[], int list[], double disp_array[],
                         std::mt19937 &rng, bool is_nvt, double &heat, double &sc_pot,
                         const boost::math::cubic_b_spline<double> &oo_spline,
                         const boost::math::cubic_b_spline<double> &oh_spline,
                         const boost::math::cubic_b_spline<double> &hh_spline,
                         const double oo_spline_cutsq, const double oh_spline_cutsq,
                         const double hh_spline_cutsq);
#endif
#endif

#ifdef QTIP4PF_ADAPTER_IMPLEMENTATION

#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

// water.cpp:29-38 (declared extern by qtip4pfparams.h in the reference)
extern double T_K, epsilonsc, dtfs, eps;
extern int n_beads, n_mols, n_atoms, n_dof, list_size;
extern std::array<double, 3> boxl;

namespace qtp_adapter {

inline qtp_ctx *&ctx_slot()
{
    static qtp_ctx *c = nullptr;
    return c;
}

inline void report(int rc, const char *where)
{
    if (rc != QTP_OK)
        std::cerr << where << ": " << qtp_last_error(ctx_slot()) << " (status " << rc << ")\n";
}

inline qtp_ctx *ctx()
{
    qtp_ctx *&c = ctx_slot();
    if (!c) {
        qtp_params p{};
        p.n_mols = n_mols;
        p.n_beads = n_beads;
        p.boxl[0] = boxl[0];
        p.boxl[1] = boxl[1];
        p.boxl[2] = boxl[2];
        p.T_K = T_K;
        p.dtfs = dtfs;
        p.eps_ewald = eps;
        p.epsilon_sc = epsilonsc;
        p.list_size = list_size;
        const int rc = qtp_create(&p, &c);
        if (rc != QTP_OK) {
            std::cerr << "qtp_create failed (status " << rc << ")"
                      << (c ? std::string(": ") + qtp_last_error(c) : std::string()) << "\n";
            std::exit(1);
        }
    }
    return c;
}

// thermostat.cpp:71, 88-98, 124 draw order
inline void draw_pile(std::mt19937 &gen, std::vector<double> &dev)
{
    std::normal_distribution<double> gauss_dev(0.0, 1.0);
    const int dof = 3 * n_atoms;
    dev.assign(3 + (size_t)(n_beads - 1) * n_dof, 0.0);
    dev[0] = gauss_dev(gen);
    if (dof % 2 == 0) {
        std::gamma_distribution<double> gamma_dev((dof - 2) / 2, 1.0);
        dev[1] = gauss_dev(gen);
        dev[2] = gamma_dev(gen);
    } else {
        std::gamma_distribution<double> gamma_dev((dof - 1) / 2, 1.0);
        dev[2] = gamma_dev(gen);
    }
    for (size_t i = 3; i < dev.size(); ++i) dev[i] = gauss_dev(gen);
}

}  // namespace qtp_adapter

double get_forces_nospline(const double position_array[], double force_array[],
                           const double charge_array[], const int point[], const int list[],
                           int bead)
{
    double e = 0.0;
    qtp_adapter::report(qtp_get_forces_nospline(qtp_adapter::ctx(), position_array, force_array,
                                                charge_array, point, list, bead, &e),
                        "get_forces_nospline");
    return e;
}

double force_eval_sc_nospline(const double position_array[], double force_array[],
                              const double charge_array[], const double mass_array[],
                              const int point[], const int list[], double &sc_pot,
                              double fphys[])
{
    double e = 0.0;
    qtp_adapter::report(qtp_force_eval_sc_nospline(qtp_adapter::ctx(), position_array,
                                                   force_array, charge_array, mass_array, point,
                                                   list, &sc_pot, fphys, &e),
                        "force_eval_sc_nospline");
    return e;
}

void neighbour_list(const double position_array[], int point[], int list[], double disp_array[],
                    int bead)
{
    const int rc = qtp_neighbour_list(qtp_adapter::ctx(), position_array, point, list, disp_array,
                                      bead);
    if (rc == QTP_ERR_LIST_OVERFLOW)
        std::cerr << "list size exceeded!\n";  // neighbour.cpp:57
    else
        qtp_adapter::report(rc, "neighbour_list");
}

void ftransform(double input_array[], const int length, const bool is_forwards)
{
    qtp_adapter::report(qtp_ftransform(qtp_adapter::ctx(), input_array, length, is_forwards ? 1 : 0),
                        "ftransform");
}

double pile_g(double momentum_array[], const double mass_array[], std::mt19937 &gen)
{
    std::vector<double> dev;
    qtp_adapter::draw_pile(gen, dev);
    double heat = 0.0;
    qtp_adapter::report(qtp_pile_g(qtp_adapter::ctx(), momentum_array, mass_array, dev.data(), &heat),
                        "pile_g");
    return heat;
}

double update_trajectory_class_nospline(double position_array[], double momentum_array[],
                                        const double mass_array[], double force_array[],
                                        double charge_array[], int point[], int list[],
                                        double disp_array[])
{
    double e = 0.0;
    qtp_adapter::report(qtp_update_trajectory_class_nospline(qtp_adapter::ctx(), position_array,
                                                             momentum_array, mass_array,
                                                             force_array, charge_array, point, list,
                                                             disp_array, &e),
                        "update_trajectory_class_nospline");
    return e;
}

double update_sc_nospline(double position_array[], double momentum_array[],
                          const double mass_array[], double force_array[], double fphys[],
                          double charge_array[], int point[], int list[], double disp_array[],
                          std::mt19937 &rng, bool is_nvt, double &heat, double &sc_pot)
{
    // pile_g is the step's only consumer of rng (update.cpp:325): drawing its
    // deviates before the step leaves the stream exactly where the reference leaves it
    std::vector<double> dev;
    if (is_nvt) qtp_adapter::draw_pile(rng, dev);
    double e = 0.0;
    qtp_adapter::report(qtp_update_sc_nospline(qtp_adapter::ctx(), position_array, momentum_array,
                                               mass_array, force_array, fphys, charge_array, point,
                                               list, disp_array, is_nvt ? dev.data() : nullptr,
                                               is_nvt ? 1 : 0, &heat, &sc_pot, &e),
                        "update_sc_nospline");
    return e;
}

double get_kcv(const double position_array[], const double force_array[])
{
    double v = 0.0;
    qtp_adapter::report(qtp_get_kcv(qtp_adapter::ctx(), position_array, force_array, &v), "get_kcv");
    return v;
}

double get_conserved(double position_array[], double momentum_array[], const double mass_array[],
                     double energy)
{
    double v = 0.0;
    qtp_adapter::report(qtp_get_conserved(qtp_adapter::ctx(), position_array, momentum_array,
                                          mass_array, energy, &v),
                        "get_conserved");
    return v;
}

#ifdef QTIP4PF_ADAPTER_SPLINES
#include <cmath>

namespace qtp_adapter {

// simulation.cpp:40-41 (and the O-H, H-H copies): every table starts at 0
// with a 0.01 bohr step; its range end is sqrt(cutsq).
constexpr double spline_start = 0.0, spline_step = 0.01;

inline void load_splines(const boost::math::cubic_b_spline<double> *sp[3], const double cutsq[3])
{
    static const void *last[3] = {nullptr, nullptr, nullptr};
    static double last_cut[3] = {-1.0, -1.0, -1.0};
    bool same = true;
    for (int t = 0; t < 3; ++t) same = same && last[t] == sp[t] && last_cut[t] == cutsq[t];
    if (same) return;
    std::vector<double> tab[3];
    for (int t = 0; t < 3; ++t) {
        const long n = std::lround((std::sqrt(cutsq[t]) - spline_start) / spline_step) + 1;
        tab[t].resize(n > 0 ? n : 0);
        for (long j = 0; j < n; ++j) tab[t][j] = (*sp[t])(spline_start + j * spline_step);
    }
    report(qtp_set_splines(ctx(), tab[0].data(), (int)tab[0].size(), tab[1].data(),
                           (int)tab[1].size(), tab[2].data(), (int)tab[2].size(), spline_start,
                           spline_step),
           "qtp_set_splines");
    for (int t = 0; t < 3; ++t) {
        last[t] = sp[t];
        last_cut[t] = cutsq[t];
    }
}

}  // namespace qtp_adapter

#define QTP_ADAPTER_LOAD_SPLINES                                                               \
    do {                                                                                      \
        const boost::math::cubic_b_spline<double> *sp_[3] = {&oo_spline, &oh_spline, &hh_spline}; \
        const double cs_[3] = {oo_spline_cutsq, oh_spline_cutsq, hh_spline_cutsq};             \
        qtp_adapter::load_splines(sp_, cs_);                                                  \
    } while (0)

double get_forces_wspline(const double position_array[], double force_array[],
                          const double charge_array[], const int point[], const int list[],
                          const boost::math::cubic_b_spline<double> &oo_spline,
                          const boost::math::cubic_b_spline<double> &oh_spline,
                          const boost::math::cubic_b_spline<double> &hh_spline,
                          const double oo_spline_cutsq, const double oh_spline_cutsq,
                          const double hh_spline_cutsq, int bead)
{
    QTP_ADAPTER_LOAD_SPLINES;
    double e = 0.0;
    qtp_adapter::report(qtp_get_forces_wspline(qtp_adapter::ctx(), position_array, force_array,
                                               charge_array, point, list, bead, &e),
                        "get_forces_wspline");
    return e;
}

double force_eval_sc_wspline(const double position_array[], double force_array[],
                             const double charge_array[], const double mass_array[],
                             const int point[], const int list[],
                             const boost::math::cubic_b_spline<double> &oo_spline,
                             const boost::math::cubic_b_spline<double> &oh_spline,
                             const boost::math::cubic_b_spline<double> &hh_spline,
                             const double oo_spline_cutsq, const double oh_spline_cutsq,
                             const double hh_spline_cutsq, double &sc_pot, double fphys[])
{
    QTP_ADAPTER_LOAD_SPLINES;
    double e = 0.0;
    qtp_adapter::report(qtp_force_eval_sc_wspline(qtp_adapter::ctx(), position_array, force_array,
                                                  charge_array, mass_array, point, list, &sc_pot,
                                                  fphys, &e),
                        "force_eval_sc_wspline");
    return e;
}

double update_trajectory_class_wspline(double position_array[], double momentum_array[],
                                       const double mass_array[], double force_array[],
                                       double charge_array[], int point[], int list[],
                                       double disp_array[],
                                       const boost::math::cubic_b_spline<double> &oo_spline,
                                       const boost::math::cubic_b_spline<double> &oh_spline,
                                       const boost::math::cubic_b_spline<double> &hh_spline,
                                       const double oo_spline_cutsq, const double oh_spline_cutsq,
                                       const double hh_spline_cutsq)
{
    QTP_ADAPTER_LOAD_SPLINES;
    double e = 0.0;
    qtp_adapter::report(qtp_update_trajectory_class_wspline(qtp_adapter::ctx(), position_array,
                                                            momentum_array, mass_array,
                                                            force_array, charge_array, point,
                                                            list, disp_array, &e),
                        "update_trajectory_class_wspline");
    return e;
}

double update_sc_wspline(double position_array[], double momentum_array[],
                         const double mass_array[], double force_array[], double fphys[],
                         double charge_array[], int point[], int list[], double disp_array[],
                         std::mt19937 &rng, bool is_nvt, double &heat, double &sc_pot,
                         const boost::math::cubic_b_spline<double> &oo_spline,
                         const boost::math::cubic_b_spline<double> &oh_spline,
                         const boost::math::cubic_b_spline<double> &hh_spline,
                         const double oo_spline_cutsq, const double oh_spline_cutsq,
                         const double hh_spline_cutsq)
{
    QTP_ADAPTER_LOAD_SPLINES;
    std::vector<double> dev;
    if (is_nvt) qtp_adapter::draw_pile(rng, dev);
    double e = 0.0;
    qtp_adapter::report(qtp_update_sc_wspline(qtp_adapter::ctx(), position_array, momentum_array,
                                              mass_array, force_array, fphys, charge_array, point,
                                              list, disp_array, is_nvt ? dev.data() : nullptr,
                                              is_nvt ? 1 : 0, &heat, &sc_pot, &e),
                        "update_sc_wspline");
    return e;
}

#undef QTP_ADAPTER_LOAD_SPLINES
#endif  // QTIP4PF_ADAPTER_SPLINES

#endif  // QTIP4PF_ADAPTER_IMPLEMENTATION
#endif  // QTIP4PF_REF_ADAPTER_HPP
