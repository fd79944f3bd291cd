// qtp_host.hpp -- host-side pieces shared by the library's translation units:
// the reference's physical constants and the host random-number generator.
#pragma once
#include <cmath>
#include <random>

struct qtp_host_rng {
    std::mt19937 gen;
};

namespace qtp {
namespace {

// qtip4pfparams.h:43-133, evaluated in the reference's order (pi truncated, :45)
struct RefConsts {
    double pi, angs2bohr, me, kb, Na, eh, kcal2au, mo, mh, K2au, au2fs, fs2au;
    double r_cut, r_cut_sq, r_skin, r_skin_sq;
    double gamma_vs, epslj, sigmalj, sigmalj6, sigmalj12, e_cut;
    double roheq, ar, ar2, ar3, ar4fac, Dr, theta_eq, ktheta, qo, qh;
    RefConsts()
    {
        pi = 3.14159265359;
        angs2bohr = 1.0 / 0.529177210903;
        me = 9.10938356e-31;
        kb = 1.38064852e-23;
        Na = 6.0221409e23;
        eh = 4.3597447222071e-18;
        kcal2au = 1.5936e-3;
        const double mhgmol = 1.0080, mogmol = 15.9994;
        mo = 0.001 * mogmol / (Na * me);
        mh = 0.001 * mhgmol / (Na * me);
        K2au = kb / eh;
        const double hbar_SI = 1.054571817e-34;
        au2fs = hbar_SI * 1e15 / eh;
        fs2au = 1.0 / au2fs;
        r_cut = 9.0 * angs2bohr;
        r_cut_sq = r_cut * r_cut;
        r_skin = 1.30 * r_cut;
        r_skin_sq = r_skin * r_skin;
        gamma_vs = 0.73612;
        epslj = 0.1852 * kcal2au;
        sigmalj = 3.1589 * angs2bohr;
        sigmalj6 = std::pow(sigmalj, 6.0);
        sigmalj12 = sigmalj6 * sigmalj6;
        e_cut = 4.0 * epslj * (std::pow((sigmalj / r_cut), 12.0) - std::pow((sigmalj / r_cut), 6.0));
        roheq = 0.9419 * angs2bohr;
        ar = 2.287 / angs2bohr;
        ar2 = ar * ar;
        ar3 = ar2 * ar;
        ar4fac = 7.0 * ar3 * ar / 12.0;
        Dr = 116.09 * kcal2au;
        theta_eq = 107.4 * pi / 180.0;
        ktheta = 87.85 * kcal2au;
        qo = -1.1128;
        qh = -0.5 * qo;
    }
};
const RefConsts RC;

}  // namespace
}  // namespace qtp
