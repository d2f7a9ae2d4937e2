// Host build of the product's per-lane device math, for the CPU tests: the
// 2- and 3-body orbit-loop polynomials (pip_orbit_loop.h) and the per-water terms
// (water_terms.h), compiled with g++ from the same headers the HIP kernels
// include.  Test infrastructure only.
#include <cmath>
using std::exp;
using std::fma;
using std::pow;
using std::sqrt;
#define __host__
#define __device__
#define __forceinline__ inline
#define __constant__
#define PIP_FENCE(v) ((void)0)
#define PIP_OPAQUE_PTR(p) ((void)0)
#define PIP_CONST_PTR(T) const T*
#define __builtin_amdgcn_sched_barrier(x) ((void)0)
#include "mbpol_params.h"
#include "pip_orbit_loop.h"
#include "water_terms.h"

extern "C" {

double host_pip2b_orbit(const double* x31, double* g31) {
    double x[31], g[31];
    for (int i = 0; i < 31; ++i) x[i] = x31[i];
    const double e = mbpol::pip2b_orbit_eval<double>(x, g);
    for (int i = 0; i < 31; ++i) g31[i] = g[i];
    return e;
}

double host_pip3b_orbit(const double* x36, double* g36) {
    double x[36], g[36];
    for (int i = 0; i < 36; ++i) x[i] = x36[i];
    const double e = mbpol::pip3b_orbit_eval<double>(x, g);
    for (int i = 0; i < 36; ++i) g36[i] = g[i];
    return e;
}

// one water O, H1, H2 (Angstrom): 1-body energy (kJ/mol) and gradient on the
// three atoms (kJ/mol/Angstrom, O H1 H2 x 3)
double host_water_onebody(const double* pos9, double* grad9) {
    mbpol::water::Geom g;
    mbpol::water::geometry(pos9, pos9 + 3, pos9 + 6, g);
    const double rHH[3] = {pos9[3] - pos9[6], pos9[4] - pos9[7], pos9[5] - pos9[8]};
    double gH1[3], gH2[3];
    const double e = mbpol::water::onebody(g, rHH, gH1, gH2);
    for (int i = 0; i < 3; ++i) {
        grad9[i] = -(gH1[i] + gH2[i]);
        grad9[3 + i] = gH1[i];
        grad9[6 + i] = gH2[i];
    }
    return e;
}

// one water O, H1, H2 (Angstrom): site charges q[4] (O, H1, H2, M) and
// dq[(atom * 3 + site) * 3 + xyz] (atom O H1 H2, site H1 H2 M; e/Angstrom)
void host_water_charges(const double* pos9, double* q4, double* dq27) {
    mbpol::water::Geom g;
    mbpol::water::geometry(pos9, pos9 + 3, pos9 + 6, g);
    mbpol::water::charges(g, q4, dq27);
}
}
