// 3-body (trimer) energy and forces -- oracle (test infrastructure only).
// Restates platforms/reference/src/MBPolReferenceThreeBodyForce.cpp:67-322.
#include "oracle_common.h"
#include "mbpol_params.h"

namespace oracle {
using namespace mbpol_params;

namespace {

// exp(-k (d - r0)) with d in Angstrom from nm positions (:67-78).
double var3(double k, double r0, const V3& a, const V3& b) {
    const V3 dx = (a - b) * nm_to_A;
    const double d = std::sqrt(dx.dot(dx));
    return std::exp(-k * (d - r0));
}

// Gradient of g * var3, already scaled to kJ/mol/nm forces (:80-101).
void g_var3(double g, double k, double r0, const V3& a, const V3& b, V3& ga, V3& gb) {
    const V3 dx = (a - b) * nm_to_A;
    const double d = std::sqrt(dx.dot(dx));
    double gg = -k * g * std::exp(-k * (d - r0)) / d;
    gg *= cal2joule * 10. * -1;
    ga += dx * gg;
    gb -= dx * gg;
}

// (1 + cos(pi (r - r3i)/(r3f - r3i)))/2 on (r3i, r3f] (:103-120).
double switch3(double r, double& g) {
    if (r > th_r3f) { g = 0.0; return 0.0; }
    if (r > th_r3i) {
        const double t1 = M_PI / (th_r3f - th_r3i);
        const double x = (r - th_r3i) * t1;
        g = -std::sin(x) * t1 / 2.0;
        return (1.0 + std::cos(x)) / 2.0;
    }
    g = 0.0;
    return 1.0;
}

struct VarDef { int a, b, kind; };  // kind: 0 HH_intra, 1 OH_intra, 2 HH, 3 OH, 4 OO

}  // namespace

// The 36 variables of poly_3b_v2x in reference order (:170-206).  Atoms are
// numbered Oa,Ha1,Ha2,Ob,Hb1,Hb2,Oc,Hc1,Hc2 = 0..8.
static const VarDef kVars3[36] = {
    {1, 2, 0}, {4, 5, 0}, {7, 8, 0},
    {0, 1, 1}, {0, 2, 1}, {3, 4, 1}, {3, 5, 1}, {6, 7, 1}, {6, 8, 1},
    {1, 4, 2}, {1, 5, 2}, {1, 7, 2}, {1, 8, 2}, {2, 4, 2}, {2, 5, 2}, {2, 7, 2}, {2, 8, 2},
    {4, 7, 2}, {4, 8, 2}, {5, 7, 2}, {5, 8, 2},
    {0, 4, 3}, {0, 5, 3}, {0, 7, 3}, {0, 8, 3}, {3, 1, 3}, {3, 2, 3}, {3, 7, 3}, {3, 8, 3},
    {6, 1, 3}, {6, 2, 3}, {6, 4, 3}, {6, 5, 3},
    {0, 3, 4}, {0, 6, 4}, {3, 6, 4}};

static void kinds3(int kind, double& k, double& r0) {
    switch (kind) {
        case 0: k = th_kHH_intra; r0 = th_dHH_intra; break;
        case 1: k = th_kOH_intra; r0 = th_dOH_intra; break;
        case 2: k = th_kHH; r0 = th_dHH; break;
        case 3: k = th_kOH; r0 = th_dOH; break;
        default: k = th_kOO; r0 = th_dOO; break;
    }
}

// One trimer (molecule indices into idx[mol][O,H1,H2]); nm in, kJ/mol out.
// MBPolReferenceThreeBodyForce.cpp:122-293.
double threebody_triplet(int si, int sj, int sk, const double* pos, const int* idx, bool periodic,
                         const double* box, double* forces) {
    const int sites[3] = {si, sj, sk};
    V3 p[9];
    for (int m = 0; m < 3; ++m)
        for (int a = 0; a < 3; ++a) p[3 * m + a] = load3(pos, idx[3 * sites[m] + a]);
    if (periodic) image_molecules(V3(box[0], box[1], box[2]), p, 3);

    const V3 rab = (p[0] - p[3]) * nm_to_A;
    const V3 rac = (p[0] - p[6]) * nm_to_A;
    const V3 rbc = (p[3] - p[6]) * nm_to_A;
    const double drab = std::sqrt(rab.dot(rab));
    const double drac = std::sqrt(rac.dot(rac));
    const double drbc = std::sqrt(rbc.dot(rbc));
    if (drab < 2 || drac < 2 || drbc < 2) return 0.;

    double x[36];
    for (int n = 0; n < 36; ++n) {
        double k, r0;
        kinds3(kVars3[n].kind, k, r0);
        x[n] = var3(k, r0, p[kVars3[n].a], p[kVars3[n].b]);
    }
    double g[36];
    double e = pip3b_eval(x, g);

    double gab, gac, gbc;
    const double sab = switch3(drab, gab);
    const double sac = switch3(drac, gac);
    const double sbc = switch3(drbc, gbc);
    const double s = sab * sac + sab * sbc + sac * sbc;
    for (int n = 0; n < 36; ++n) g[n] *= s;

    V3 f[9];
    for (int n = 0; n < 36; ++n) {
        double k, r0;
        kinds3(kVars3[n].kind, k, r0);
        g_var3(g[n], k, r0, p[kVars3[n].a], p[kVars3[n].b], f[kVars3[n].a], f[kVars3[n].b]);
    }
    gab *= (sac + sbc) * e / drab;
    gac *= (sab + sbc) * e / drac;
    gbc *= (sab + sac) * e / drbc;
    e *= s;
    for (int n = 0; n < 3; ++n) {
        f[0][n] += (gab * rab[n] + gac * rac[n]) * cal2joule * -nm_to_A;
        f[3][n] += (gbc * rbc[n] - gab * rab[n]) * cal2joule * -nm_to_A;
        f[6][n] -= (gac * rac[n] + gbc * rbc[n]) * cal2joule * -nm_to_A;
    }
    for (int m = 0; m < 3; ++m)
        for (int a = 0; a < 3; ++a) add3(forces, idx[3 * sites[m] + a], f[3 * m + a]);
    return e * cal2joule;
}

// Loop over a trimer list (MBPolReferenceThreeBodyForce.cpp:295-322).
double threebody_all(int64_t ntrip, const int* trip, const double* pos, const int* idx, bool periodic,
                     const double* box, double* forces) {
    double e = 0.0;
    for (int64_t t = 0; t < ntrip; ++t)
        e += threebody_triplet(trip[3 * t], trip[3 * t + 1], trip[3 * t + 2], pos, idx, periodic, box, forces);
    return e;
}

}  // namespace oracle
