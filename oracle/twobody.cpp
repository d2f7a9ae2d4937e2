// 2-body (dimer) energy and forces -- oracle (test infrastructure only).
// Restates platforms/reference/src/MBPolReferenceTwoBodyForce.cpp:110-323 and
// the helpers of mbpol_2body_constants.cpp:10-111.
#include "oracle_common.h"
#include "mbpol_params.h"

namespace oracle {
using namespace mbpol_params;

namespace {

// A distance variable with its gradient direction kept for the chain rule
// (mbpol_2body_constants.cpp:10-58).
struct DistVar {
    V3 g;
    double v_exp(double r0, double k, const V3& a, const V3& b) {
        g = a - b;
        const double r = std::sqrt(g.dot(g));
        const double e = std::exp(k * (r0 - r));
        g *= -k * e / r;
        return e;
    }
    double v_coul(double r0, double k, const V3& a, const V3& b) {
        g = a - b;
        const double r = std::sqrt(g.dot(g));
        const double e = std::exp(k * (r0 - r));
        const double ri = 1.0 / r;
        const double val = e * ri;
        g *= -(k + ri) * val * ri;
        return val;
    }
    void grads(double gg, V3& a, V3& b) const {
        V3 d = g * gg;
        a += d;
        b -= d;
    }
};

// Extra points X1, X2 of one monomer and their back-propagation
// (mbpol_2body_constants.cpp:62-96).
struct Monomer {
    V3 oh1, oh2;
    void setup(const V3& O, const V3& H1, const V3& H2, V3& x1, V3& x2) {
        oh1 = H1 - O;
        oh2 = H2 - O;
        const V3 v = oh1.cross(oh2);
        const V3 in_plane = O + (oh1 + oh2) * (0.5 * tb_in_plane_gamma);
        const V3 out_of_plane = v * tb_out_of_plane_gamma;
        x1 = in_plane + out_of_plane;
        x2 = in_plane - out_of_plane;
    }
    void grads(const V3& g1, const V3& g2, V3& O, V3& H1, V3& H2) const {
        const V3 gm = g1 - g2;
        const V3 t1 = oh2.cross(gm);
        const V3 t2 = oh1.cross(gm);
        const V3 gsum = g1 + g2;
        const V3 in_plane = gsum * (0.5 * tb_in_plane_gamma);
        const V3 gh1 = in_plane + t1 * tb_out_of_plane_gamma;
        const V3 gh2 = in_plane - t2 * tb_out_of_plane_gamma;
        O += gsum - (gh1 + gh2);
        H1 += gh1;
        H2 += gh2;
    }
};

// Cosine switch from r2i to r2f (mbpol_2body_constants.cpp:98-111).
double switch2(double r, double& g) {
    if (r > tb_r2f) { g = 0.0; return 0.0; }
    if (r > tb_r2i) {
        const double t1 = M_PI / (tb_r2f - tb_r2i);
        const double x = (r - tb_r2i) * t1;
        g = -std::sin(x) * t1 / 2.0;
        return (1.0 + std::cos(x)) / 2.0;
    }
    g = 0.0;
    return 1.0;
}

}  // namespace

// One dimer; site indices into the O-only molecule table `idx` ([mol][O,H1,H2]).
// Positions in nm; forces accumulated in kJ/mol/nm; returns kJ/mol.
// MBPolReferenceTwoBodyForce.cpp:110-296.
double twobody_pair(int si, int sj, const double* pos, const int* idx, bool periodic, const double* box,
                    double* forces) {
    enum { Oa, Ha1, Ha2, Ob, Hb1, Hb2 };
    enum { Xa1, Xa2, Xb1, Xb2 };
    V3 p[6];
    for (int a = 0; a < 3; ++a) p[a] = load3(pos, idx[3 * si + a]) * nm_to_A;
    for (int a = 0; a < 3; ++a) p[3 + a] = load3(pos, idx[3 * sj + a]) * nm_to_A;
    if (periodic) image_molecules(V3(box[0], box[1], box[2]) * nm_to_A, p, 2);

    const V3 dOO = p[Oa] - p[Ob];
    const double rOO = std::sqrt(dOO.dot(dOO));
    if (rOO > tb_r2f) return 0.0;
    if (rOO < 2.) return 0.0;

    Monomer ma, mb;
    V3 xp[4];
    ma.setup(p[Oa], p[Ha1], p[Ha2], xp[Xa1], xp[Xa2]);
    mb.setup(p[Ob], p[Hb1], p[Hb2], xp[Xb1], xp[Xb2]);

    const double d0i = tb_d0_intra, d0e = tb_d0_inter;
    DistVar c[31];
    double v[31];
    v[0] = c[0].v_exp(d0i, tb_k_HH_intra, p[Ha1], p[Ha2]);
    v[1] = c[1].v_exp(d0i, tb_k_HH_intra, p[Hb1], p[Hb2]);
    v[2] = c[2].v_exp(d0i, tb_k_OH_intra, p[Oa], p[Ha1]);
    v[3] = c[3].v_exp(d0i, tb_k_OH_intra, p[Oa], p[Ha2]);
    v[4] = c[4].v_exp(d0i, tb_k_OH_intra, p[Ob], p[Hb1]);
    v[5] = c[5].v_exp(d0i, tb_k_OH_intra, p[Ob], p[Hb2]);
    v[6] = c[6].v_coul(d0e, tb_k_HH_coul, p[Ha1], p[Hb1]);
    v[7] = c[7].v_coul(d0e, tb_k_HH_coul, p[Ha1], p[Hb2]);
    v[8] = c[8].v_coul(d0e, tb_k_HH_coul, p[Ha2], p[Hb1]);
    v[9] = c[9].v_coul(d0e, tb_k_HH_coul, p[Ha2], p[Hb2]);
    v[10] = c[10].v_coul(d0e, tb_k_OH_coul, p[Oa], p[Hb1]);
    v[11] = c[11].v_coul(d0e, tb_k_OH_coul, p[Oa], p[Hb2]);
    v[12] = c[12].v_coul(d0e, tb_k_OH_coul, p[Ob], p[Ha1]);
    v[13] = c[13].v_coul(d0e, tb_k_OH_coul, p[Ob], p[Ha2]);
    v[14] = c[14].v_coul(d0e, tb_k_OO_coul, p[Oa], p[Ob]);
    v[15] = c[15].v_exp(d0e, tb_k_XH_main, xp[Xa1], p[Hb1]);
    v[16] = c[16].v_exp(d0e, tb_k_XH_main, xp[Xa1], p[Hb2]);
    v[17] = c[17].v_exp(d0e, tb_k_XH_main, xp[Xa2], p[Hb1]);
    v[18] = c[18].v_exp(d0e, tb_k_XH_main, xp[Xa2], p[Hb2]);
    v[19] = c[19].v_exp(d0e, tb_k_XH_main, xp[Xb1], p[Ha1]);
    v[20] = c[20].v_exp(d0e, tb_k_XH_main, xp[Xb1], p[Ha2]);
    v[21] = c[21].v_exp(d0e, tb_k_XH_main, xp[Xb2], p[Ha1]);
    v[22] = c[22].v_exp(d0e, tb_k_XH_main, xp[Xb2], p[Ha2]);
    v[23] = c[23].v_exp(d0e, tb_k_XO_main, p[Oa], xp[Xb1]);
    v[24] = c[24].v_exp(d0e, tb_k_XO_main, p[Oa], xp[Xb2]);
    v[25] = c[25].v_exp(d0e, tb_k_XO_main, p[Ob], xp[Xa1]);
    v[26] = c[26].v_exp(d0e, tb_k_XO_main, p[Ob], xp[Xa2]);
    v[27] = c[27].v_exp(d0e, tb_k_XX_main, xp[Xa1], xp[Xb1]);
    v[28] = c[28].v_exp(d0e, tb_k_XX_main, xp[Xa1], xp[Xb2]);
    v[29] = c[29].v_exp(d0e, tb_k_XX_main, xp[Xa2], xp[Xb1]);
    v[30] = c[30].v_exp(d0e, tb_k_XX_main, xp[Xa2], xp[Xb2]);

    double g[31];
    const double E_poly = pip2b_eval(v, g);

    V3 f[6], fx[4];
    c[0].grads(g[0], f[Ha1], f[Ha2]);
    c[1].grads(g[1], f[Hb1], f[Hb2]);
    c[2].grads(g[2], f[Oa], f[Ha1]);
    c[3].grads(g[3], f[Oa], f[Ha2]);
    c[4].grads(g[4], f[Ob], f[Hb1]);
    c[5].grads(g[5], f[Ob], f[Hb2]);
    c[6].grads(g[6], f[Ha1], f[Hb1]);
    c[7].grads(g[7], f[Ha1], f[Hb2]);
    c[8].grads(g[8], f[Ha2], f[Hb1]);
    c[9].grads(g[9], f[Ha2], f[Hb2]);
    c[10].grads(g[10], f[Oa], f[Hb1]);
    c[11].grads(g[11], f[Oa], f[Hb2]);
    c[12].grads(g[12], f[Ob], f[Ha1]);
    c[13].grads(g[13], f[Ob], f[Ha2]);
    c[14].grads(g[14], f[Oa], f[Ob]);
    c[15].grads(g[15], fx[Xa1], f[Hb1]);
    c[16].grads(g[16], fx[Xa1], f[Hb2]);
    c[17].grads(g[17], fx[Xa2], f[Hb1]);
    c[18].grads(g[18], fx[Xa2], f[Hb2]);
    c[19].grads(g[19], fx[Xb1], f[Ha1]);
    c[20].grads(g[20], fx[Xb1], f[Ha2]);
    c[21].grads(g[21], fx[Xb2], f[Ha1]);
    c[22].grads(g[22], fx[Xb2], f[Ha2]);
    c[23].grads(g[23], f[Oa], fx[Xb1]);
    c[24].grads(g[24], f[Oa], fx[Xb2]);
    c[25].grads(g[25], f[Ob], fx[Xa1]);
    c[26].grads(g[26], f[Ob], fx[Xa2]);
    c[27].grads(g[27], fx[Xa1], fx[Xb1]);
    c[28].grads(g[28], fx[Xa1], fx[Xb2]);
    c[29].grads(g[29], fx[Xa2], fx[Xb1]);
    c[30].grads(g[30], fx[Xa2], fx[Xb2]);
    ma.grads(fx[Xa1], fx[Xa2], f[Oa], f[Ha1], f[Ha2]);
    mb.grads(fx[Xb1], fx[Xb2], f[Ob], f[Hb1], f[Hb2]);

    double gsw;
    const double sw = switch2(rOO, gsw);
    const double s = sw * cal2joule * -10.;
    for (int a = 0; a < 3; ++a) {
        add3(forces, idx[3 * si + a], f[a] * s);
        add3(forces, idx[3 * sj + a], f[3 + a] * s);
    }
    gsw *= E_poly / rOO;
    for (int i = 0; i < 3; ++i) {
        const double d = gsw * dOO[i];
        forces[3 * idx[3 * si] + i] += d * cal2joule * -10.;
        forces[3 * idx[3 * sj] + i] -= d * cal2joule * -10.;
    }
    return sw * E_poly * cal2joule;
}

// Loop over a dimer list (MBPolReferenceTwoBodyForce.cpp:298-323).
double twobody_all(int npairs, const int* pairs, const double* pos, const int* idx, bool periodic,
                   const double* box, double* forces) {
    double e = 0.0;
    for (int p = 0; p < npairs; ++p) e += twobody_pair(pairs[2 * p], pairs[2 * p + 1], pos, idx, periodic, box, forces);
    return e;
}

}  // namespace oracle
