// 1-body (monomer) energy and forces -- oracle (test infrastructure only).
// Restates platforms/reference/src/MBPolReferenceOneBodyForce.cpp:69-221.
#include "oracle_common.h"
#include "mbpol_params.h"

namespace oracle {
using namespace mbpol_params;

// One water, positions in nm; forces accumulated in kJ/mol/nm; returns kJ/mol.
// MBPolReferenceOneBodyForce.cpp:69-201.
double onebody_water(const V3& pO, const V3& pH1, const V3& pH2, V3& fO, V3& fH1, V3& fH2) {
    double r1[3], r2[3], rhh[3], d1 = 0, d2 = 0, dhh = 0;
    for (int i = 0; i < 3; ++i) {
        r1[i] = (pH1[i] - pO[i]) * 10.;
        r2[i] = (pH2[i] - pO[i]) * 10.;
        rhh[i] = (pH1[i] - pH2[i]) * 10.;
        d1 += r1[i] * r1[i];
        d2 += r2[i] * r2[i];
        dhh += rhh[i] * rhh[i];
    }
    d1 = std::sqrt(d1);
    d2 = std::sqrt(d2);
    dhh = std::sqrt(dhh);
    const double costh = (r1[0] * r2[0] + r1[1] * r2[1] + r1[2] * r2[2]) / (d1 * d2);

    const double phh1 = ob_f5z * ob_phh1A * std::exp(ob_phh2);
    const double e1 = std::exp(-ob_alphaoh * (d1 - ob_roh));
    const double e2 = std::exp(-ob_alphaoh * (d2 - ob_roh));
    const double Va = ob_deoh * (e1 * (e1 - 2.0) + e2 * (e2 - 2.0));
    const double Vb = phh1 * std::exp(-ob_phh2 * dhh);
    const double dVa1 = 2.0 * ob_alphaoh * ob_deoh * e1 * (1.0 - e1) / d1;
    const double dVa2 = 2.0 * ob_alphaoh * ob_deoh * e2 * (1.0 - e2) / d2;
    const double dVb = -ob_phh2 * Vb / dhh;

    const double x1 = (d1 - ob_reoh) / ob_reoh;
    const double x2 = (d2 - ob_reoh) / ob_reoh;
    const double x3 = costh - ob_costhe;
    // powers x^(j-1) with pw[.][0] = 0 so that (j-1)*x^(j-2) vanishes for j=1
    double pw[3][16];
    for (int i = 0; i < 3; ++i) { pw[i][0] = 0.0; pw[i][1] = 1.0; }
    for (int j = 2; j < 16; ++j) {
        pw[0][j] = pw[0][j - 1] * x1;
        pw[1][j] = pw[1][j - 1] * x2;
        pw[2][j] = pw[2][j - 1] * x3;
    }
    const double efac = std::exp(-ob_b1 * (std::pow(d1 - ob_reoh, 2) + std::pow(d2 - ob_reoh, 2)));
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int j = 1; j < 245; ++j) {
        const int a = ob_idx1[j], b = ob_idx2[j], c = ob_idx3[j];
        const double cz = ob_c5z[j];
        s0 += cz * (pw[0][a] * pw[1][b] + pw[0][b] * pw[1][a]) * pw[2][c];
        s1 += cz * ((a - 1) * pw[0][a - 1] * pw[1][b] + (b - 1) * pw[0][b - 1] * pw[1][a]) * pw[2][c];
        s2 += cz * ((b - 1) * pw[0][a] * pw[1][b - 1] + (a - 1) * pw[0][b] * pw[1][a - 1]) * pw[2][c];
        s3 += cz * (pw[0][a] * pw[1][b] + pw[0][b] * pw[1][a]) * (c - 1) * pw[2][c - 1];
    }
    const double Vc = 2 * ob_c5z[0] + efac * s0;
    double e = Va + Vb + Vc;
    e += ob_e1_shift;
    e *= ob_cm1_kcalmol;

    const double dVcdr1 = (-2 * ob_b1 * efac * (d1 - ob_reoh) * s0 + efac * s1 / ob_reoh) / d1;
    const double dVcdr2 = (-2 * ob_b1 * efac * (d2 - ob_reoh) * s0 + efac * s2 / ob_reoh) / d2;
    const double dVcdcth = efac * s3;
    for (int i = 0; i < 3; ++i) {
        const double gH1 = (dVa1 * r1[i] + dVb * rhh[i] + dVcdr1 * r1[i] +
                            dVcdcth * (r2[i] / (d1 * d2) - costh * r1[i] / (d1 * d1))) *
                           ob_cm1_kcalmol * cal2joule * 10;
        const double gH2 = (dVa2 * r2[i] - dVb * rhh[i] + dVcdr2 * r2[i] +
                            dVcdcth * (r1[i] / (d1 * d2) - costh * r2[i] / (d2 * d2))) *
                           ob_cm1_kcalmol * cal2joule * 10;
        fH1[i] -= gH1;
        fH2[i] -= gH2;
        fO[i] -= -(gH1 + gH2);
    }
    return e * cal2joule;
}

// MBPolReferenceOneBodyForce.cpp:204-221 (periodic: hydrogens imaged to their O).
double onebody_all(int nw, const double* pos, const int* idx, bool periodic, const double* box, double* forces) {
    double energy = 0.0;
    V3 b = box ? V3(box[0], box[1], box[2]) : V3();
    for (int w = 0; w < nw; ++w) {
        V3 p[3];
        for (int a = 0; a < 3; ++a) p[a] = load3(pos, idx[3 * w + a]);
        if (periodic) image_molecules(b, p, 1);
        V3 f[3];
        energy += onebody_water(p[0], p[1], p[2], f[0], f[1], f[2]);
        for (int a = 0; a < 3; ++a) add3(forces, idx[3 * w + a], f[a]);
    }
    return energy;
}

}  // namespace oracle
