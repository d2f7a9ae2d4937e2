// Per-water terms of MB-pol, shared by the device kernel and a host build
// (the CPU test compiles this header with g++ against the oracle).
//
// One water, positions in Angstrom.  Both per-water surfaces are functions of
// the same internal coordinates -- the OH bond lengths d1, d2 and
// c = cos(HOH) -- so the geometry and the chain rule from (d1, d2, c) to the
// atoms are computed once and serve both:
//   * the 1-body energy (platforms/reference/src/MBPolReferenceOneBodyForce.cpp:69-201):
//       V = D[e1(e1 - 2) + e2(e2 - 2)]  + p_HH exp(-k_HH r_HH)  + 2 c0 + f(d1, d2) S(x),
//       e_i = exp(-a (d_i - r_OH)),  f = exp(-b ((d1 - r_e)^2 + (d2 - r_e)^2)),
//     in cm^-1, shifted and converted to kJ/mol;
//   * the dipole-moment surface (MBPolElectrostaticsForce.cpp:2793-3096):
//       p1 = d0 + f_D T(x1, x2, x3) + B L,  p2 = d0 + f_D T(x2, x1, x3) + B L,
//       L = a (d1^b + d2^b)(c0 + c1 c + c2 P2(c)),  B = Bohr radius in Angstrom,
//     and the site charges q_H1 = p1 + g (p1 + p2), q_H2 = p2 + g (p1 + p2),
//     q_M = -(p1 + p2)/(1 - gamma_M), q_O = 0, with g = (gamma_M / 2)/(1 - gamma_M).
// S and T are generated polynomials (water_poly_gen.h) in
// x1 = (d1 - r_e)/r_e, x2 = (d2 - r_e)/r_e, x3 = c - cos(theta_e).
#pragma once
#include <math.h>

#include "mbpol_params.h"
#include "water_poly_gen.h"

namespace mbpol {
namespace water {
using namespace mbpol_params;

struct Geom {
    double r1[3], r2[3];  // H1 - O, H2 - O (Angstrom)
    double d1, d2, c;     // bond lengths and cos(HOH)
};

__host__ __device__ __forceinline__ void geometry(const double O[3], const double H1[3], const double H2[3], Geom& g) {
    for (int i = 0; i < 3; ++i) {
        g.r1[i] = H1[i] - O[i];
        g.r2[i] = H2[i] - O[i];
    }
    g.d1 = sqrt(g.r1[0] * g.r1[0] + g.r1[1] * g.r1[1] + g.r1[2] * g.r1[2]);
    g.d2 = sqrt(g.r2[0] * g.r2[0] + g.r2[1] * g.r2[1] + g.r2[2] * g.r2[2]);
    g.c = (g.r1[0] * g.r2[0] + g.r1[1] * g.r2[1] + g.r1[2] * g.r2[2]) / (g.d1 * g.d2);
}

// Gradient of a function of (d1, d2, c) on H1 and H2 from its partial
// derivatives (the O gradient is minus their sum):
//   d d1/d r1 = r1/d1,   d c/d r1 = r2/(d1 d2) - c r1/d1^2   (and 1 <-> 2).
__host__ __device__ __forceinline__ void to_atoms(const Geom& g, double dd1, double dd2, double dc, double gH1[3],
                                                  double gH2[3]) {
    const double i12 = 1.0 / (g.d1 * g.d2);
    const double a1 = (dd1 - dc * g.c / g.d1) / g.d1, a2 = (dd2 - dc * g.c / g.d2) / g.d2, b = dc * i12;
    for (int i = 0; i < 3; ++i) {
        gH1[i] = a1 * g.r1[i] + b * g.r2[i];
        gH2[i] = a2 * g.r2[i] + b * g.r1[i];
    }
}

// 1-body energy (kJ/mol) and its gradient on H1 and H2 (kJ/mol/Angstrom; the
// O gradient is minus their sum).
__host__ __device__ __forceinline__ double onebody(const Geom& g, const double rHH[3], double gH1[3], double gH2[3]) {
    const double re = ob_reoh;
    const double x1 = (g.d1 - re) / re, x2 = (g.d2 - re) / re, x3 = g.c - ob_costhe;
    double S1, S2, S3;
    const double S = onebody_poly(x1, x2, x3, S1, S2, S3);
    const double u1 = g.d1 - re, u2 = g.d2 - re;
    const double f = exp(-ob_b1 * (u1 * u1 + u2 * u2));
    // Morse stretches
    const double e1 = exp(-ob_alphaoh * (g.d1 - ob_roh)), e2 = exp(-ob_alphaoh * (g.d2 - ob_roh));
    const double vstretch = ob_deoh * (e1 * (e1 - 2.0) + e2 * (e2 - 2.0));
    // H-H repulsion
    const double dHH = sqrt(rHH[0] * rHH[0] + rHH[1] * rHH[1] + rHH[2] * rHH[2]);
    const double pHH = ob_f5z * ob_phh1A * exp(ob_phh2);
    const double vHH = pHH * exp(-ob_phh2 * dHH);
    const double V = vstretch + vHH + 2.0 * onebody_poly_c0 + f * S;
    // partial derivatives (cm^-1 per Angstrom, per unit c)
    const double dd1 = 2.0 * ob_alphaoh * ob_deoh * e1 * (1.0 - e1) + f * (S1 / re - 2.0 * ob_b1 * u1 * S);
    const double dd2 = 2.0 * ob_alphaoh * ob_deoh * e2 * (1.0 - e2) + f * (S2 / re - 2.0 * ob_b1 * u2 * S);
    const double dc = f * S3;
    const double unit = ob_cm1_kcalmol * cal2joule;
    to_atoms(g, dd1 * unit, dd2 * unit, dc * unit, gH1, gH2);
    const double hh = -ob_phh2 * vHH / dHH * unit;
    for (int i = 0; i < 3; ++i) {
        gH1[i] += hh * rHH[i];
        gH2[i] -= hh * rHH[i];
    }
    return (V + ob_e1_shift) * unit;
}

// Site charges q[4] (O, H1, H2, M; e) and their derivatives
// dq[(atom * 3 + site) * 3 + xyz] for atom O, H1, H2 and site H1, H2, M (e/Angstrom).
__host__ __device__ __forceinline__ void charges(const Geom& g, double q[4], double dq[27]) {
    const double re = dms_reoh;
    const double x1 = (g.d1 - re) / re, x2 = (g.d2 - re) / re, x3 = g.c - dms_costhe;
    double Ta, Tb, Tc, Ua, Ub, Uc;
    const double T = dipole_poly(x1, x2, x3, Ta, Tb, Tc);  // p1 (H1 first)
    const double U = dipole_poly(x2, x1, x3, Ua, Ub, Uc);  // p2 (H2 first): Ua is d/dx2, Ub d/dx1
    const double u1 = g.d1 - re, u2 = g.d2 - re;
    const double f = exp(-dms_b1D * (u1 * u1 + u2 * u2));
    // long-range part B L(d1, d2, c), common to both
    const double B = dms_Bohr_A;
    const double P2 = 0.5 * (3.0 * g.c * g.c - 1.0);
    const double ang = dms_c0 + dms_c1 * g.c + dms_c2 * P2;
    const double pw1 = pow(g.d1, dms_b), pw2 = pow(g.d2, dms_b);
    const double L = dms_a * (pw1 + pw2) * ang;
    const double Ld1 = dms_a * dms_b * pw1 / g.d1 * ang, Ld2 = dms_a * dms_b * pw2 / g.d2 * ang;
    const double Lc = dms_a * (pw1 + pw2) * (dms_c1 + 3.0 * dms_c2 * g.c);
    const double p1 = dipole_poly_d0 + f * T + B * L;
    const double p2 = dipole_poly_d0 + f * U + B * L;
    const double f1 = -2.0 * dms_b1D * u1 * f, f2 = -2.0 * dms_b1D * u2 * f;  // df/dd1, df/dd2
    double g1H1[3], g1H2[3], g2H1[3], g2H2[3];
    to_atoms(g, f * Ta / re + f1 * T + B * Ld1, f * Tb / re + f2 * T + B * Ld2, f * Tc + B * Lc, g1H1, g1H2);
    to_atoms(g, f * Ub / re + f1 * U + B * Ld1, f * Ua / re + f2 * U + B * Ld2, f * Uc + B * Lc, g2H1, g2H2);
    const double gm = 1.0 - dms_gammaM, gs = 0.5 * dms_gammaM / gm;
    q[0] = 0.0;
    q[1] = p1 + gs * (p1 + p2);
    q[2] = p2 + gs * (p1 + p2);
    q[3] = -(p1 + p2) / gm;
    for (int i = 0; i < 3; ++i) {
        const double a1[3] = {-(g1H1[i] + g1H2[i]), g1H1[i], g1H2[i]};  // dp1 / d(O, H1, H2)
        const double a2[3] = {-(g2H1[i] + g2H2[i]), g2H1[i], g2H2[i]};  // dp2 / d(O, H1, H2)
        for (int at = 0; at < 3; ++at) {
            const double sum = a1[at] + a2[at];
            dq[(at * 3 + 0) * 3 + i] = a1[at] + gs * sum;
            dq[(at * 3 + 1) * 3 + i] = a2[at] + gs * sum;
            dq[(at * 3 + 2) * 3 + i] = -sum / gm;
        }
    }
}

}  // namespace water
}  // namespace mbpol
