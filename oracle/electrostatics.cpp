// Thole-damped, polarisable MB-pol electrostatics (direct sum and PME) --
// oracle (test infrastructure only).
//
// Restates platforms/reference/src/MBPolReferenceElectrostaticsForce.cpp.
// The reference carries two induced-dipole sets ("d" and "p", AMOEBA
// heritage).  They start equal and are updated identically, and this file
// keeps both so that every formula reads as the reference's does.
#include <algorithm>
#include <complex>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <chrono>

#include "oracle_common.h"
#include "mbpol_params.h"
#include "oracle.h"

namespace oracle {
using namespace mbpol_params;
typedef std::complex<double> cplx;
void fft3d(cplx* grid, int nx, int ny, int nz, int sign);

// ---------------------------------------------------------------------------
// Incomplete gamma function Q(a, x) = Gamma(a, x) / Gamma(a): power series for
// x < a + 1 and a modified-Lentz continued fraction otherwise (the method of
// Press et al., Numerical Recipes 3rd ed., sec. 6.2, which the reference's
// gammq.cpp:70-176 also follows).
// ---------------------------------------------------------------------------
double gammq(double a, double x) {
    const double eps = std::numeric_limits<double>::epsilon();
    const double fpmin = std::numeric_limits<double>::min() / eps;
    if (x == 0.0) return 1.0;
    const double gln = std::lgamma(a);
    if (x < a + 1.0) {
        double ap = a, sum = 1.0 / a, del = sum;
        for (;;) {
            ap += 1.0;
            del *= x / ap;
            sum += del;
            if (std::fabs(del) < std::fabs(sum) * eps) return 1.0 - sum * std::exp(-x + a * std::log(x) - gln);
        }
    }
    double b = x + 1.0 - a, c = 1.0 / fpmin, d = 1.0 / b, h = d;
    for (int i = 1;; ++i) {
        const double an = -i * (i - a);
        b += 2.0;
        d = an * d + b;
        if (std::fabs(d) < fpmin) d = fpmin;
        c = b + an / c;
        if (std::fabs(c) < fpmin) c = fpmin;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (std::fabs(del - 1.0) <= eps) break;
    }
    return std::exp(-x + a * std::log(x) - gln) * h;
}

namespace {

enum { TCC = 0, TCD = 1, TDD = 2, TDDOH = 3, TDDHH = 4 };

struct Site {
    V3 pos;
    double q, damp, pol;
    int mol, type;
    V3 dq[3];       // d q_(site s of this water) / d r_(this site), per nm
    int other[3];   // indices of the H1, H2, M sites of this water
};

class Elec {
   public:
    Elec(const oracle_elec_params& P, int n) : P_(P), n_(n) {
        for (int i = 0; i < 5; ++i) thole_[i] = P.thole[i];
        box_ = V3(P.box[0], P.box[1], P.box[2]);
        pme_ = P.method == 1;
        expgamm_ = std::tgamma(0.75);
    }

    // getAndScaleInverseRs (:260-328)
    double scale(const Site& a, const Site& b, double r, bool just_scale, int order, int itype) const {
        const bool same = a.mol == b.mol;
        double rr = 1.;
        if (!just_scale) {
            const double ri = 1.0 / r, r2i = ri * ri;
            rr = ri;
            for (int k = 3; k <= order; k += 2) rr *= (k - 2) * r2i;
        }
        double pg = thole_[itype];
        if (itype == TDD) {
            if (same) pg = (a.type == 0 || b.type == 0) ? thole_[TDDOH] : thole_[TDDHH];
            else pg = thole_[TDD];
        }
        const double damp = std::pow(a.damp * b.damp, 1 / 6.);
        if (damp != 0.0 && damp > -50.0) {
            const double ratio = std::pow(r / damp, 4);
            const double dfe = -1 * pg * ratio;
            switch (order) {
                case 1:
                    return rr * (1.0 - std::exp(dfe) + std::pow(pg, 1.0 / 4.0) * (r / damp) * expgamm_ * gammq(3.0 / 4.0, -dfe));
                case 3:
                    return rr * (1.0 - std::exp(dfe));
                case 5:
                    return rr * ((1.0 - std::exp(dfe)) - (4. / 3.) * pg * std::exp(dfe) * ratio);
                case 7:
                    return rr * (((1.0 - std::exp(dfe)) - (4. / 3.) * pg * std::exp(dfe) * ratio) -
                                 (4. / 15.) * pg * (4. * pg * ratio - 1.) * std::exp(dfe) / std::pow(damp, 4) * std::pow(r, 4));
            }
        }
        return rr;
    }

    // getAndScaleInverseRs13justScaleTCC (:330-359)
    void scale13_tcc(const Site& a, const Site& b, double pg, double r, double& s1, double& s3) const {
        const double damp = std::pow(a.damp * b.damp, 1 / 6.);
        if (damp != 0.0 && damp > -50.0) {
            const double ratio = std::pow(r / damp, 4);
            const double dfe = -1 * pg * ratio;
            s3 = 1.0 - std::exp(dfe);
            s1 = s3 + std::pow(pg, 1.0 / 4.0) * (r / damp) * expgamm_ * gammq(3.0 / 4.0, -dfe);
        } else {
            s3 = 1.;
            s1 = 1.;
        }
    }

    // getPeriodicDelta (:1234-1239); no-op without PME (header :834)
    void periodic_delta(V3& d) const {
        if (!pme_) return;
        for (int k = 0; k < 3; ++k) d[k] -= std::floor(d[k] * (1.0 / box_[k]) + 0.5) * box_[k];
    }

    // computeWaterCharge (:2793-3096)
    static void water_charge(Site& O, Site& H1, Site& H2, Site& M, int iO, int iH1, int iH2, int iM) {
        const double gamma1 = 1.0 - dms_gammaM;
        const double gamma2 = dms_gammaM / 2;
        const double xx = dms_Bohr_A, xx2 = xx * xx;
        double R1[3], R2[3], d1 = 0, d2 = 0;
        for (int i = 0; i < 3; ++i) {
            R1[i] = H1.pos[i] * 10. - O.pos[i] * 10.;
            R2[i] = H2.pos[i] * 10. - O.pos[i] * 10.;
            d1 += R1[i] * R1[i];
            d2 += R2[i] * R2[i];
        }
        d1 = std::sqrt(d1);
        d2 = std::sqrt(d2);
        const double costh = (R1[0] * R2[0] + R1[1] * R2[1] + R1[2] * R2[2]) / (d1 * d2);
        const double efac = std::exp(-dms_b1D * (std::pow(d1 - dms_reoh, 2) + std::pow(d2 - dms_reoh, 2)));
        const double x1 = (d1 - dms_reoh) / dms_reoh, x2 = (d2 - dms_reoh) / dms_reoh, x3 = costh - dms_costhe;
        double pw[3][16];
        for (int i = 0; i < 3; ++i) { pw[i][0] = 0.0; pw[i][1] = 1.0; }
        for (int j = 2; j < 16; ++j) {
            pw[0][j] = pw[0][j - 1] * x1;
            pw[1][j] = pw[1][j - 1] * x2;
            pw[2][j] = pw[2][j - 1] * x3;
        }
        double p1 = 0, p2 = 0, dp1dr1 = 0, dp1dr2 = 0, dp1dc = 0, dp2dr1 = 0, dp2dr2 = 0, dp2dc = 0;
        const double pl1 = costh, pl2 = 0.5 * (3 * pl1 * pl1 - 1.0);
        for (int j = 1; j < 84; ++j) {
            const int a = dms_idx0[j], b = dms_idx1[j], c = dms_idx2[j];
            const double cf = dms_coef[j];
            p1 += cf * pw[0][a] * pw[1][b] * pw[2][c];
            p2 += cf * pw[0][b] * pw[1][a] * pw[2][c];
            dp1dr1 += cf * (a - 1) * pw[0][a - 1] * pw[1][b] * pw[2][c];
            dp1dr2 += cf * (b - 1) * pw[0][a] * pw[1][b - 1] * pw[2][c];
            dp1dc += cf * (c - 1) * pw[0][a] * pw[1][b] * pw[2][c - 1];
            dp2dr1 += cf * (b - 1) * pw[0][b - 1] * pw[1][a] * pw[2][c];
            dp2dr2 += cf * (a - 1) * pw[0][b] * pw[1][a - 1] * pw[2][c];
            dp2dc += cf * (c - 1) * pw[0][b] * pw[1][a] * pw[2][c - 1];
        }
        dp1dr1 /= dms_reoh / xx;
        dp1dr2 /= dms_reoh / xx;
        dp2dr1 /= dms_reoh / xx;
        dp2dr2 /= dms_reoh / xx;
        const double pc0 = dms_a * (std::pow(d1, dms_b) + std::pow(d2, dms_b)) * (dms_c0 + pl1 * dms_c1 + pl2 * dms_c2);
        const double dpc0dr1 = dms_a * dms_b * std::pow(d1, dms_b - 1) * (dms_c0 + pl1 * dms_c1 + pl2 * dms_c2) * xx2;
        const double dpc0dr2 = dms_a * dms_b * std::pow(d2, dms_b - 1) * (dms_c0 + pl1 * dms_c1 + pl2 * dms_c2) * xx2;
        const double dpc0dc = dms_a * (std::pow(d1, dms_b) + std::pow(d2, dms_b)) * (dms_c1 + 0.5 * (6.0 * pl1) * dms_c2) * xx;
        const double defdr1 = -2.0 * dms_b1D * (d1 - dms_reoh) * efac * xx;
        const double defdr2 = -2.0 * dms_b1D * (d2 - dms_reoh) * efac * xx;
        dp1dr1 = dp1dr1 * efac + p1 * defdr1 + dpc0dr1;
        dp1dr2 = dp1dr2 * efac + p1 * defdr2 + dpc0dr2;
        dp1dc = dp1dc * efac + dpc0dc;
        dp2dr1 = dp2dr1 * efac + p2 * defdr1 + dpc0dr1;
        dp2dr2 = dp2dr2 * efac + p2 * defdr2 + dpc0dr2;
        dp2dc = dp2dc * efac + dpc0dc;
        p1 = dms_coef[0] + p1 * efac + pc0 * xx;
        p2 = dms_coef[0] + p2 * efac + pc0 * xx;
        const double qO = -(p1 + p2), g21 = gamma2 / gamma1;
        O.q = 0.;
        H1.q = p1 + g21 * (p1 + p2);
        H2.q = p2 + g21 * (p1 + p2);
        M.q = qO / gamma1;
        dp1dr1 /= xx;
        dp1dr2 /= xx;
        dp2dr1 /= xx;
        dp2dr2 /= xx;
        const double f1q1r13 = (dp1dr1 - (dp1dc * costh / d1)) / d1;
        const double f1q1r23 = dp1dc / (d1 * d2);
        const double f2q1r23 = (dp1dr2 - (dp1dc * costh / d2)) / d2;
        const double f2q1r13 = dp1dc / (d2 * d1);
        const double f1q2r13 = (dp2dr1 - (dp2dc * costh / d1)) / d1;
        const double f1q2r23 = dp2dc / (d1 * d2);
        const double f2q2r23 = (dp2dr2 - (dp2dc * costh / d2)) / d2;
        const double f2q2r13 = dp2dc / (d2 * d1);
        // dq1/dr, dq2/dr w.r.t. H1, H2, O (second index: charge differentiated)
        double q1H1[3], q1H2[3], q1O[3], q2H1[3], q2H2[3], q2O[3], qOH1[3], qOH2[3], qOO[3];
        for (int i = 0; i < 3; ++i) {
            q1H1[i] = f1q1r13 * R1[i] + f1q1r23 * R2[i];
            q1H2[i] = f2q1r13 * R1[i] + f2q1r23 * R2[i];
            q1O[i] = -(q1H1[i] + q1H2[i]);
            q2H1[i] = f1q2r13 * R1[i] + f1q2r23 * R2[i];
            q2H2[i] = f2q2r13 * R1[i] + f2q2r23 * R2[i];
            q2O[i] = -(q2H1[i] + q2H2[i]);
            qOH1[i] = -(q1H1[i] + q2H1[i]);
            qOH2[i] = -(q1H2[i] + q2H2[i]);
            qOO[i] = -(q1O[i] + q2O[i]);
        }
        for (int i = 0; i < 3; ++i) {
            for (int s = 0; s < 3; ++s) M.dq[s][i] = 0.;
            const double sH1 = g21 * (q1H1[i] + q2H1[i]);
            const double sH2 = g21 * (q1H2[i] + q2H2[i]);
            const double sO = g21 * (q1O[i] + q2O[i]);
            H1.dq[0][i] = q1H1[i] + sH1;
            H2.dq[0][i] = q1H2[i] + sH2;
            O.dq[0][i] = q1O[i] + sO;
            H1.dq[1][i] = q2H1[i] + sH1;
            H2.dq[1][i] = q2H2[i] + sH2;
            O.dq[1][i] = q2O[i] + sO;
            H1.dq[2][i] = qOH1[i] - 2 * sH1;
            H2.dq[2][i] = qOH2[i] - 2 * sH2;
            O.dq[2][i] = qOO[i] - 2 * sO;
        }
        for (int s = 0; s < 3; ++s) {
            H1.dq[s] *= 10;
            H2.dq[s] *= 10;
            O.dq[s] *= 10;
        }
        Site* all[4] = {&O, &H1, &H2, &M};
        for (Site* st : all) {
            st->other[0] = iH1;
            st->other[1] = iH2;
            st->other[2] = iM;
        }
        (void)iO;
    }

    // ---------------- PME pieces (:1098-2111) ----------------
    void pme_setup() {
        for (int d = 0; d < 3; ++d) ng_[d] = P_.grid[d];
        ntot_ = (size_t)ng_[0] * ng_[1] * ng_[2];
        grid_.assign(ntot_, cplx(0, 0));
        // B-spline moduli (:1248-1340)
        const int order = el_pme_order;
        double arr[8];
        {
            const double x = 0.0;
            arr[0] = 1.0 - x;
            arr[1] = x;
            for (int k = 2; k < order; ++k) {
                const double den = 1.0 / k;
                arr[k] = x * arr[k - 1] * den;
                for (int i = 1; i < k; ++i) arr[k - i] = ((x + i) * arr[k - i - 1] + ((k - i + 1) - x) * arr[k - i]) * den;
                arr[0] = (1.0 - x) * arr[0] * den;
            }
        }
        for (int dim = 0; dim < 3; ++dim) {
            const int n = ng_[dim];
            std::vector<double> bs(std::max(n, order + 2) + 1, 0.0);
            for (int i = 2; i <= order + 1; ++i) bs[i] = arr[i - 2];
            std::vector<double>& mod = moduli_[dim];
            mod.assign(n, 0.0);
            const double factor = 2.0 * M_PI / n;
            for (int i = 0; i < n; ++i) {
                double s1 = 0, s2 = 0;
                for (int j = 1; j <= n; ++j) {
                    const double ang = factor * i * (j - 1);
                    s1 += bs[j] * std::cos(ang);
                    s2 += bs[j] * std::sin(ang);
                }
                mod[i] = s1 * s1 + s2 * s2;
            }
            const double tiny = 1.0e-7;
            if (mod[0] < tiny) mod[0] = 0.5 * mod[1];
            for (int i = 1; i < n - 1; ++i)
                if (mod[i] < tiny) mod[i] = 0.5 * (mod[i - 1] + mod[i + 1]);
            if (mod[n - 1] < tiny) mod[n - 1] = 0.5 * mod[n - 2];
            for (int i = 1; i <= n; ++i) {
                int k = i - 1;
                if (i > n / 2) k = k - n;
                double zeta;
                if (k == 0) {
                    zeta = 1.0;
                } else {
                    double s1 = 1.0, s2 = 1.0;
                    const double f = M_PI * k / n;
                    for (int j = 1; j <= 50; ++j) {
                        const double a = f / (f + M_PI * j);
                        s1 = s1 + std::pow(a, order);
                        s2 = s2 + std::pow(a, 2 * order);
                    }
                    for (int j = 1; j <= 50; ++j) {
                        const double a = f / (f - M_PI * j);
                        s1 += std::pow(a, order);
                        s2 += std::pow(a, 2 * order);
                    }
                    zeta = s2 / s1;
                }
                mod[i - 1] = mod[i - 1] * (zeta * zeta);
            }
        }
    }

    // computeBSplinePoint (:1446-1519): theta[i] = {value, d1, d2, d3}
    static void bspline(double w, double theta[5][4]) {
        const int N = 5;
        double A[6][6];  // A[order][i], 1-based
        A[2][2] = w;
        A[2][1] = 1.0 - w;
        A[3][3] = 0.5 * w * A[2][2];
        A[3][2] = 0.5 * ((1.0 + w) * A[2][1] + (2.0 - w) * A[2][2]);
        A[3][1] = 0.5 * (1.0 - w) * A[2][1];
        for (int i = 4; i <= N; ++i) {
            const int k = i - 1;
            const double den = 1.0 / k;
            A[i][i] = den * w * A[k][k];
            for (int j = 1; j <= i - 2; ++j) A[i][i - j] = den * ((w + j) * A[k][i - j - 1] + (i - j - w) * A[k][i - j]);
            A[i][1] = den * (1.0 - w) * A[k][1];
        }
        // first derivative coefficients
        int k = N - 1;
        A[k][N] = A[k][N - 1];
        for (int i = N - 1; i >= 2; --i) A[k][i] = A[k][i - 1] - A[k][i];
        A[k][1] = -A[k][1];
        // second derivative
        k = N - 2;
        A[k][N - 1] = A[k][N - 2];
        for (int i = N - 2; i >= 2; --i) A[k][i] = A[k][i - 1] - A[k][i];
        A[k][1] = -A[k][1];
        A[k][N] = A[k][N - 1];
        for (int i = N - 1; i >= 2; --i) A[k][i] = A[k][i - 1] - A[k][i];
        A[k][1] = -A[k][1];
        // third derivative
        k = N - 3;
        A[k][N - 2] = A[k][N - 3];
        for (int i = N - 3; i >= 2; --i) A[k][i] = A[k][i - 1] - A[k][i];
        A[k][1] = -A[k][1];
        A[k][N - 1] = A[k][N - 2];
        for (int i = N - 2; i >= 2; --i) A[k][i] = A[k][i - 1] - A[k][i];
        A[k][1] = -A[k][1];
        A[k][N] = A[k][N - 1];
        for (int i = N - 1; i >= 2; --i) A[k][i] = A[k][i - 1] - A[k][i];
        A[k][1] = -A[k][1];
        for (int i = 1; i <= N; ++i) {
            theta[i - 1][0] = A[N][i];
            theta[i - 1][1] = A[N - 1][i];
            theta[i - 1][2] = A[N - 2][i];
            theta[i - 1][3] = A[N - 3][i];
        }
    }

    // computeMBPolBsplines (:1524-1554)
    void pme_splines(const std::vector<Site>& s) {
        theta_.assign((size_t)n_ * 3 * 5 * 4, 0.0);
        igrid_.assign((size_t)n_ * 3, 0);
        for (int i = 0; i < n_; ++i) {
            V3 p = s[i].pos;
            periodic_delta(p);
            for (int d = 0; d < 3; ++d) {
                const double w0 = p[d] * (1.0 / box_[d]);
                const double fr = ng_[d] * (w0 - (int)(w0 + 0.5) + 0.5);
                const int ifr = (int)fr;
                const double w = fr - ifr;
                int g = ifr - el_pme_order + 1;
                g += g < 0 ? ng_[d] : 0;
                igrid_[3 * i + d] = g;
                double th[5][4];
                bspline(w, th);
                for (int k = 0; k < 5; ++k)
                    for (int q = 0; q < 4; ++q) theta_[(((size_t)i * 3 + d) * 5 + k) * 4 + q] = th[k][q];
            }
        }
    }
    double th(int i, int d, int k, int q) const { return theta_[(((size_t)i * 3 + d) * 5 + k) * 4 + q]; }
    size_t gidx(int x, int y, int z) const { return ((size_t)x * ng_[1] + y) * ng_[2] + z; }
    int wrap(int v, int d) const { return v >= ng_[d] ? v - ng_[d] : v; }

    // spread charges (real) / fractional dipoles d (real) and p (imag)
    // (:1599-1674, :1823-1904) -- written as a scatter, same sums.
    void spread_charges(const std::vector<Site>& s) {
        std::fill(grid_.begin(), grid_.end(), cplx(0, 0));
        for (int i = 0; i < n_; ++i)
            for (int ix = 0; ix < 5; ++ix)
                for (int iy = 0; iy < 5; ++iy)
                    for (int iz = 0; iz < 5; ++iz) {
                        const size_t g = gidx(wrap(igrid_[3 * i] + ix, 0), wrap(igrid_[3 * i + 1] + iy, 1), wrap(igrid_[3 * i + 2] + iz, 2));
                        grid_[g] += cplx(s[i].q * th(i, 1, iy, 0) * th(i, 2, iz, 0) * th(i, 0, ix, 0), 0.0);
                    }
    }
    void spread_dipoles(const std::vector<V3>& mud, const std::vector<V3>& mup) {
        std::fill(grid_.begin(), grid_.end(), cplx(0, 0));
        V3 sc(ng_[0] / box_[0], ng_[1] / box_[1], ng_[2] / box_[2]);
        for (int i = 0; i < n_; ++i) {
            V3 d(sc[0] * mud[i][0], sc[1] * mud[i][1], sc[2] * mud[i][2]);
            V3 p(sc[0] * mup[i][0], sc[1] * mup[i][1], sc[2] * mup[i][2]);
            for (int ix = 0; ix < 5; ++ix)
                for (int iy = 0; iy < 5; ++iy)
                    for (int iz = 0; iz < 5; ++iz) {
                        const double u0 = th(i, 1, iy, 0), u1 = th(i, 1, iy, 1), v0 = th(i, 2, iz, 0), v1 = th(i, 2, iz, 1);
                        const double t0 = th(i, 0, ix, 0), t1 = th(i, 0, ix, 1);
                        const double re = (d[1] * u1 * v0 + d[2] * u0 * v1) * t0 + d[0] * u0 * v0 * t1;
                        const double im = (p[1] * u1 * v0 + p[2] * u0 * v1) * t0 + p[0] * u0 * v0 * t1;
                        const size_t g = gidx(wrap(igrid_[3 * i] + ix, 0), wrap(igrid_[3 * i + 1] + iy, 1), wrap(igrid_[3 * i + 2] + iz, 2));
                        grid_[g] += cplx(re, im);
                    }
        }
    }

    // performMBPolReciprocalConvolution (:1676-1713)
    void convolve() {
        const double expf = (M_PI * M_PI) / (P_.alpha * P_.alpha);
        const double sf = 1.0 / (M_PI * box_[0] * box_[1] * box_[2]);
        for (int kx = 0; kx < ng_[0]; ++kx)
            for (int ky = 0; ky < ng_[1]; ++ky)
                for (int kz = 0; kz < ng_[2]; ++kz) {
                    const size_t g = gidx(kx, ky, kz);
                    if (kx == 0 && ky == 0 && kz == 0) { grid_[g] = 0; continue; }
                    const int mx = (kx < (ng_[0] + 1) / 2) ? kx : (kx - ng_[0]);
                    const int my = (ky < (ng_[1] + 1) / 2) ? ky : (ky - ng_[1]);
                    const int mz = (kz < (ng_[2] + 1) / 2) ? kz : (kz - ng_[2]);
                    const double hx = mx * (1.0 / box_[0]), hy = my * (1.0 / box_[1]), hz = mz * (1.0 / box_[2]);
                    const double m2 = hx * hx + hy * hy + hz * hz;
                    const double den = m2 * moduli_[0][kx] * moduli_[1][ky] * moduli_[2][kz];
                    const double et = sf * std::exp(-expf * m2) / den;
                    grid_[g] *= et;
                }
    }

    void fft_roundtrip() {
        fft3d(grid_.data(), ng_[0], ng_[1], ng_[2], -1);
        convolve();
        fft3d(grid_.data(), ng_[0], ng_[1], ng_[2], +1);
    }

    // gather the potential and its derivatives up to second order, for the real
    // and imaginary grid parts (:1715-1821, :1906-2111).  Component order:
    // 0:000 1:100 2:010 3:001 4:200 5:020 6:002 7:110 8:101 9:011.
    void gather(int i, double re[10], double im[10]) const {
        for (int c = 0; c < 10; ++c) re[c] = im[c] = 0;
        for (int iz = 0; iz < 5; ++iz) {
            const int k = wrap(igrid_[3 * i + 2] + iz, 2);
            double v[4];
            for (int q = 0; q < 4; ++q) v[q] = th(i, 2, iz, q);
            for (int iy = 0; iy < 5; ++iy) {
                const int j = wrap(igrid_[3 * i + 1] + iy, 1);
                double u[4];
                for (int q = 0; q < 4; ++q) u[q] = th(i, 1, iy, q);
                for (int ix = 0; ix < 5; ++ix) {
                    const int ii = wrap(igrid_[3 * i] + ix, 0);
                    const cplx gv = grid_[gidx(ii, j, k)];
                    double t[4];
                    for (int q = 0; q < 4; ++q) t[q] = th(i, 0, ix, q);
                    const double w[10] = {t[0] * u[0] * v[0], t[1] * u[0] * v[0], t[0] * u[1] * v[0], t[0] * u[0] * v[1],
                                          t[2] * u[0] * v[0], t[0] * u[2] * v[0], t[0] * u[0] * v[2], t[1] * u[1] * v[0],
                                          t[1] * u[0] * v[1], t[0] * u[1] * v[1]};
                    for (int c = 0; c < 10; ++c) {
                        re[c] += gv.real() * w[c];
                        im[c] += gv.imag() * w[c];
                    }
                }
            }
        }
    }

    // ---------------- the full calculation ----------------
    double run(std::vector<Site>& s, double* forces, oracle_elec_stats* stats, double* dipoles_out, double* fixed_field_out) {
        const auto t_start = std::chrono::steady_clock::now();
        // setup (:857-897): charges from the dipole-moment surface
        if (P_.charge_redistribution) {
            for (int ii = 0; ii + 3 < n_; ii += 4) water_charge(s[ii], s[ii + 1], s[ii + 2], s[ii + 3], ii, ii + 1, ii + 2, ii + 3);
        }
        if (pme_) {
            pme_setup();
            pme_splines(s);
        }
        // ---- fixed field (:402-420 direct; :1409-1439 PME) ----
        std::vector<V3> Ed(n_), Ep(n_);
        if (pme_) {
            spread_charges(s);
            fft_roundtrip();
            phi_fixed_.assign((size_t)n_ * 10, 0.0);
            for (int i = 0; i < n_; ++i) {
                double re[10], im[10];
                gather(i, re, im);
                for (int c = 0; c < 10; ++c) phi_fixed_[10 * i + c] = re[c];
                for (int d = 0; d < 3; ++d) Ed[i][d] = -(ng_[d] / box_[d]) * re[1 + d];
                Ep[i] = Ed[i];
            }
        }
        for (int i = 0; i < n_; ++i)
            for (int j = i + 1; j < n_; ++j) {
                if (!pme_) {
                    if (s[i].mol == s[j].mol) continue;
                    const V3 dr = s[j].pos - s[i].pos;
                    const double r = std::sqrt(dr.dot(dr));
                    const double rr3 = scale(s[i], s[j], r, false, 3, TCC);
                    const V3 fi = dr * (rr3 * s[j].q), fj = dr * (rr3 * s[i].q);
                    Ed[i] -= fi; Ep[i] -= fi;
                    Ed[j] += fj; Ep[j] += fj;
                } else {
                    V3 dr = s[j].pos - s[i].pos;
                    periodic_delta(dr);
                    const double r2 = dr.dot(dr);
                    if (r2 > P_.cutoff * P_.cutoff) continue;
                    const double r = std::sqrt(r2);
                    const double ra = P_.alpha * r;
                    const double bn0 = std::erfc(ra) / r;
                    double alsq2n = 1.0 / (el_sqrt_pi * P_.alpha);
                    const double ex = std::exp(-(ra * ra));
                    alsq2n *= 2.0 * P_.alpha * P_.alpha;
                    const double bn1 = (bn0 + alsq2n * ex) / r2;
                    double s3 = scale(s[i], s[j], r, true, 3, TCC);
                    if (s[i].mol == s[j].mol) s3 = 2;
                    const double rr3 = (s3 - 1.) / (r2 * r);
                    const V3 fim = dr * (-bn1 * s[j].q), fjm = dr * (bn1 * s[i].q);
                    const V3 fid = dr * (-rr3 * s[j].q), fjd = dr * (rr3 * s[i].q);
                    Ed[i] += fim - fid; Ed[j] += fjm - fjd;
                    Ep[i] += fim - fid; Ep[j] += fjm - fjd;
                }
            }
        if (fixed_field_out)
            for (int i = 0; i < n_; ++i)
                for (int d = 0; d < 3; ++d) fixed_field_out[3 * i + d] = Ed[i][d];
        for (int i = 0; i < n_; ++i) { Ed[i] *= s[i].pol; Ep[i] *= s[i].pol; }

        // ---- induced dipoles (:618-647, :557-616) ----
        std::vector<V3> mud = Ed, mup = Ep;
        int iters = 0;
        double eps = 0;
        const auto t_fixed = std::chrono::steady_clock::now();
        bool converged = solve_dipoles(s, Ed, Ep, mud, mup, iters, eps);
        if (stats) {
            stats->iterations = iters;
            stats->epsilon = eps;
            stats->converged = converged ? 1 : 0;
            stats->setup_s = std::chrono::duration<double>(t_fixed - t_start).count() + scale_seconds_;
            stats->iteration_s = iters > 0 ? loop_seconds_ / iters : 0.0;
        }
        if (dipoles_out)
            for (int i = 0; i < n_; ++i)
                for (int d = 0; d < 3; ++d) dipoles_out[3 * i + d] = mud[i][d];
        if (!converged) throw std::runtime_error("Induced dipoles did not converge: iterations=" + std::to_string(iters) +
                                                 " eps=" + std::to_string(eps));
        const auto t_energy = std::chrono::steady_clock::now();
        const double e = pme_ ? energy_pme(s, mud, mup, forces) : energy_direct(s, mud, mup, forces);
        if (stats) stats->energy_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_energy).count();
        return e;
    }

    // SOR iteration with the reference's stop rules (:557-616, :488-531).
    bool solve_dipoles(const std::vector<Site>& s, const std::vector<V3>& Ed, const std::vector<V3>& Ep, std::vector<V3>& mud,
                       std::vector<V3>& mup, int& iters, double& eps_out) {
        // precomputeScale35 (:534-555 direct, :2280-2298 PME); pairs beyond the
        // PME cutoff are never used, so only in-cutoff pairs are kept.
        const auto t_scales = std::chrono::steady_clock::now();
        struct PairS { int i, j; V3 dr; double s3, s5; };
        std::vector<PairS> pairs;
        for (int i = 0; i < n_; ++i)
            for (int j = i + 1; j < n_; ++j) {
                V3 dr = s[j].pos - s[i].pos;
                periodic_delta(dr);
                const double r2 = dr.dot(dr);
                const double r = std::sqrt(r2);
                if (pme_) {
                    if (r2 > P_.cutoff * P_.cutoff) continue;
                    pairs.push_back({i, j, dr, scale(s[i], s[j], r, true, 3, TDD), scale(s[i], s[j], r, true, 5, TDD)});
                } else {
                    pairs.push_back({i, j, dr, -1 * scale(s[i], s[j], r, false, 3, TDD), scale(s[i], s[j], r, false, 5, TDD)});
                }
            }
        const auto t_loop = std::chrono::steady_clock::now();
        scale_seconds_ = std::chrono::duration<double>(t_loop - t_scales).count();
        double cur = 1.0e+50;
        bool done = false, conv = false;
        int it = 0;
        std::vector<V3> fd(n_), fp(n_);
        while (!done) {
            std::fill(fd.begin(), fd.end(), V3());
            std::fill(fp.begin(), fp.end(), V3());
            for (const PairS& pr : pairs) {
                double p1, p2;
                if (pme_) {
                    const double r2 = pr.dr.dot(pr.dr), r = std::sqrt(r2);
                    const double ra = P_.alpha * r;
                    const double bn0 = std::erfc(ra) / r;
                    double alsq2n = 1.0 / (el_sqrt_pi * P_.alpha);
                    const double ex = std::exp(-(ra * ra));
                    const double alsq2 = 2.0 * P_.alpha * P_.alpha;
                    alsq2n *= alsq2;
                    const double bn1 = (bn0 + alsq2n * ex) / r2;
                    alsq2n *= alsq2;
                    const double bn2 = (3.0 * bn1 + alsq2n * ex) / r2;
                    const double r3 = r * r2, r5 = r3 * r2;
                    const double rr3 = (1.0 - pr.s3) / r3, rr5 = 3.0 * (1.0 - pr.s5) / r5;
                    p1 = rr3 - bn1;
                    p2 = bn2 - rr5;
                } else {
                    p1 = pr.s3;
                    p2 = pr.s5;
                }
                // calculateInducedDipolePairIxn (:438-486) / PME (:2374-2390)
                fd[pr.i] += pr.dr * (mud[pr.j].dot(pr.dr) * p2) + mud[pr.j] * p1;
                fd[pr.j] += pr.dr * (mud[pr.i].dot(pr.dr) * p2) + mud[pr.i] * p1;
                fp[pr.i] += pr.dr * (mup[pr.j].dot(pr.dr) * p2) + mup[pr.j] * p1;
                fp[pr.j] += pr.dr * (mup[pr.i].dot(pr.dr) * p2) + mup[pr.i] * p1;
            }
            if (pme_) {
                // calculateReciprocalSpaceInducedDipoleField (:2326-2337) + self (:2357-2364)
                spread_dipoles(mud, mup);
                fft_roundtrip();
                phid_.assign((size_t)n_ * 10, 0.0);
                phip_.assign((size_t)n_ * 10, 0.0);
                for (int i = 0; i < n_; ++i) {
                    double re[10], im[10];
                    gather(i, re, im);
                    for (int c = 0; c < 10; ++c) { phid_[10 * i + c] = re[c]; phip_[10 * i + c] = im[c]; }
                    for (int d = 0; d < 3; ++d) {
                        fd[i][d] += -(ng_[d] / box_[d]) * re[1 + d];
                        fp[i][d] += -(ng_[d] / box_[d]) * im[1 + d];
                    }
                }
                const double term = (4.0 / 3.0) * (P_.alpha * P_.alpha * P_.alpha) / el_sqrt_pi;
                for (int i = 0; i < n_; ++i) { fd[i] += mud[i] * term; fp[i] += mup[i] * term; }
            }
            // updateInducedDipole (:516-531) for both sets
            double e1 = 0, e2 = 0;
            for (int i = 0; i < n_; ++i) {
                V3 nd = Ed[i] + fd[i] * s[i].pol;
                V3 dd = nd - mud[i];
                mud[i] = mud[i] + dd * el_polar_sor;
                e1 += dd.dot(dd);
                V3 np = Ep[i] + fp[i] * s[i].pol;
                V3 dp = np - mup[i];
                mup[i] = mup[i] + dp * el_polar_sor;
                e2 += dp.dot(dp);
            }
            double e = std::max(e1, e2);
            e = el_polar_sor * el_debye * std::sqrt(e / (double)n_);
            if (e < P_.target_eps) {
                conv = true;
                done = true;
            } else if (cur < e || it >= P_.max_iter) {
                done = true;
            }
            cur = e;
            ++it;
        }
        iters = it;
        eps_out = cur;
        loop_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_loop).count();
        return conv;
    }

    // calculateElectrostaticPairIxn + loop (:649-855), direct sum.
    double energy_direct(const std::vector<Site>& s, const std::vector<V3>& mud, const std::vector<V3>& mup, double* F) {
        const double f = el_electric;
        double energy = 0.0;
        for (int ii = 0; ii < n_; ++ii)
            for (int kk = ii + 1; kk < n_; ++kk) {
                const Site &I = s[ii], &K = s[kk];
                const V3 dl = K.pos - I.pos;
                const double r2 = dl.dot(dl), r = std::sqrt(r2);
                const double rr1 = 1.0 / r, rr3 = rr1 / r2, rr5 = 3.0 * rr3 / r2, rr7 = 5.0 * rr5 / r2;
                const double sci1 = mud[ii].dot(mud[kk]);
                const double sci2 = mud[ii].dot(dl), sci3 = mud[kk].dot(dl);
                const double scip1 = mud[ii].dot(mup[kk]) + mup[ii].dot(mud[kk]);
                const double scip2 = mup[ii].dot(dl), scip3 = mup[kk].dot(dl);
                (void)sci1;
                double gl0 = I.q * K.q;
                double gli0 = K.q * sci2 - I.q * sci3;
                double glip0 = K.q * scip2 - I.q * scip3;
                const bool same = I.mol == K.mol;
                if (same) gl0 = gli0 = glip0 = 0.;
                const double s1CC = scale(I, K, r, true, 1, TCC), s3CD = scale(I, K, r, true, 3, TCD);
                double e = rr1 * gl0 * s1CC;
                e += 0.5 * (rr3 * gli0 * s3CD);
                e *= f;
                const double s3CC = scale(I, K, r, true, 3, TCC), s5CD = scale(I, K, r, true, 5, TCD);
                const double s5DD = scale(I, K, r, true, 5, TDD), s7DD = scale(I, K, r, true, 7, TDD);
                const double gf0 = rr3 * gl0 * s3CC;
                const double gfi0 = 0.5 * rr5 * gli0 * s5CD + 0.5 * rr5 * glip0 * s5CD + 0.5 * rr5 * scip1 * s5DD -
                                    0.5 * rr7 * (sci2 * scip3 + scip2 * sci3) * s7DD;
                V3 ftm2 = dl * gf0;
                V3 ftm2i = dl * gfi0;
                ftm2i += (mup[ii] * sci3 + mud[ii] * scip3 + mup[kk] * sci2 + mud[kk] * scip2) * (0.5 * rr5 * s5DD);
                if (!same) ftm2i += ((mud[ii] + mup[ii]) * -K.q + (mud[kk] + mup[kk]) * I.q) * (0.5 * rr3 * s3CD);
                if (P_.charge_redistribution && !same) {
                    const double pg = thole_[TCC];
                    for (int q = 0; q < 3; ++q) {
                        const Site& oi = s[I.other[q]];
                        const Site& ok = s[K.other[q]];
                        const V3 dI = oi.pos - K.pos, dK = ok.pos - I.pos;
                        const double rI = std::sqrt(dI.dot(dI)), rK = std::sqrt(dK.dot(dK));
                        double s1I, s3I, s1K, s3K;
                        scale13_tcc(oi, K, pg, rI, s1I, s3I);
                        scale13_tcc(ok, I, pg, rK, s1K, s3K);
                        const double uI = mud[kk].dot(dI), uK = mud[ii].dot(dK);
                        for (int c = 0; c < 3; ++c) {
                            ftm2[c] += s1I * (1.0 / rI) * I.dq[q][c] * K.q;
                            ftm2[c] -= s1K * (1.0 / rK) * K.dq[q][c] * I.q;
                            ftm2i[c] += s3I * std::pow(1.0 / rI, 3) * I.dq[q][c] * uI;
                            ftm2i[c] -= s3K * std::pow(1.0 / rK, 3) * K.dq[q][c] * uK;
                        }
                    }
                }
                V3 force = (ftm2 + ftm2i) * f;
                add3(F, ii, force * -1.0);
                add3(F, kk, force);
                energy += e;
            }
        return energy;
    }

    // PME energy/forces (:2510-2777).
    double energy_pme(const std::vector<Site>& s, const std::vector<V3>& mud, const std::vector<V3>& mup, double* F) {
        const double f = el_electric;
        std::vector<double> V(n_, 0.0);
        double energy = 0.0;
        // direct space (:2510-2716)
        for (int ii = 0; ii < n_; ++ii)
            for (int jj = ii + 1; jj < n_; ++jj) {
                const Site &I = s[ii], &J = s[jj];
                V3 dr = J.pos - I.pos;
                periodic_delta(dr);
                const double r2 = dr.dot(dr);
                if (r2 > P_.cutoff * P_.cutoff) continue;
                const double r = std::sqrt(r2), ck = J.q, ci = I.q;
                const double ra = P_.alpha * r;
                const double bn0 = std::erfc(ra) / r;
                const double alsq2 = 2.0 * P_.alpha * P_.alpha;
                double alsq2n = P_.alpha > 0.0 ? 1.0 / (el_sqrt_pi * P_.alpha) : 0.0;
                const double ex = std::exp(-(ra * ra));
                alsq2n *= alsq2;
                const double bn1 = (bn0 + alsq2n * ex) / r2;
                alsq2n *= alsq2;
                const double bn2 = (3.0 * bn1 + alsq2n * ex) / r2;
                alsq2n *= alsq2;
                const double bn3 = (5.0 * bn2 + alsq2n * ex) / r2;
                const double rr1 = 1.0 / r, rr3 = rr1 / r2, rr5 = 3.0 * rr3 / r2, rr7 = 5.0 * rr5 / r2;
                const double sci3 = mud[ii].dot(dr), sci4 = mud[jj].dot(dr);
                const double scip2 = mud[ii].dot(mup[jj]) + mup[ii].dot(mud[jj]);
                const double scip3 = mup[ii].dot(dr), scip4 = mup[jj].dot(dr);
                const double gl0 = ci * ck;
                const double gli1 = ck * sci3 - ci * sci4;
                const double glip1 = ck * scip3 - ci * scip4;
                const bool same = I.mol == J.mol;
                double e = bn0 * gl0, ei = 0.5 * (bn1 * gli1);
                double s1CC = scale(I, J, r, true, 1, TCC), s3CD = scale(I, J, r, true, 3, TCD);
                if (same) s1CC = s3CD = 0.;
                const double erl = rr1 * gl0 * (1 - s1CC);
                const double erli = 0.5 * (rr3 * gli1 * (1 - s3CD));
                e = e - erl;
                ei = ei - erli;
                double energyp = (e + ei);
                V[ii] += ck * (bn0 - rr1 * (1 - s1CC));
                V[jj] += ci * (bn0 - rr1 * (1 - s1CC));
                V[ii] -= sci4 * (bn1 - rr3 * (1 - s3CD));
                V[jj] += sci3 * (bn1 - rr3 * (1 - s3CD));
                double s3CC = 0., s5CD = 0.;
                const double s5DD = scale(I, J, r, true, 5, TDD), s7DD = scale(I, J, r, true, 7, TDD);
                if (!same) {
                    s3CC = scale(I, J, r, true, 3, TCC);
                    s5CD = scale(I, J, r, true, 5, TCD);
                }
                const double gf1 = bn1 * gl0, gfr1 = (1 - s3CC) * rr3 * gl0;
                const double gfi1 = 0.5 * (bn2 * (gli1 + glip1 + scip2) - bn3 * (sci3 * scip4 + scip3 * sci4));
                const double gfi2 = -ck * bn1, gfi3 = ci * bn1;
                const double gfri1 = 0.5 * (rr5 * (gli1 * (1 - s5CD) + glip1 * (1 - s5CD) + scip2 * (1 - s5DD)) -
                                            rr7 * (sci3 * scip4 + scip3 * sci4) * (1 - s7DD));
                V3 ftm2 = dr * gf1, ftm2r = dr * gfr1, ftm2i = dr * gfi1;
                ftm2i += ((mud[ii] + mup[ii]) * gfi2 + (mud[jj] + mup[jj]) * gfi3 + (mup[ii] * sci4 + mud[ii] * scip4) * bn2 +
                          (mup[jj] * sci3 + mud[jj] * scip3) * bn2) * 0.5;
                V3 ftm2ri = dr * gfri1;
                ftm2ri += (mup[ii] * sci4 + mud[ii] * scip4 + mup[jj] * sci3 + mud[jj] * scip3) * (0.5 * rr5 * (1 - s5DD));
                ftm2ri += ((mud[ii] + mup[ii]) * -ck + (mud[jj] + mup[jj]) * ci) * (0.5 * rr3 * (1 - s3CD));
                ftm2 -= ftm2r;
                ftm2i -= ftm2ri;
                energyp *= f;
                add3(F, ii, (ftm2 + ftm2i) * -f);
                add3(F, jj, (ftm2 + ftm2i) * f);
                energy += energyp;
            }
        const V3 sc(ng_[0] / box_[0], ng_[1] / box_[1], ng_[2] / box_[2]);
        // reciprocal, induced dipoles (:2186-2265); the induced potentials are the
        // ones left by the last solver iteration.  The reference multiplies
        // phidp components 4..9 by multipole[1..9], which it never initialises
        // (:2189, :2244-2248); MB-pol sites carry charges only, so those terms
        // are taken as zero here.
        static const int d1[] = {1, 4, 7, 8}, d2[] = {2, 7, 5, 9}, d3[] = {3, 8, 9, 6};
        double erind = 0.0;
        for (int i = 0; i < n_; ++i) {
            const double* phi = &phi_fixed_[10 * i];
            const double* pd = &phid_[10 * i];
            const double* pp = &phip_[10 * i];
            double pdp[10];
            for (int c = 0; c < 10; ++c) pdp[c] = pd[c] + pp[c];
            // phidp[0] is the full induced potential (phid[0] is stored as 0)
            erind += sc[0] * mud[i][0] * phi[1] + sc[1] * mud[i][1] * phi[2] + sc[2] * mud[i][2] * phi[3];
            V[i] += .5 * pdp[0];
            double fv[3] = {0, 0, 0};
            for (int k = 0; k < 3; ++k) {
                const double u = mud[i][k] + mup[i][k];
                fv[0] += u * phi[d1[k + 1]] * (sc[k] / sc[0]);
                fv[1] += u * phi[d2[k + 1]] * (sc[k] / sc[1]);
                fv[2] += u * phi[d3[k + 1]] * (sc[k] / sc[2]);
                fv[0] += (mud[i][k] * pp[d1[k + 1]] + mup[i][k] * pd[d1[k + 1]]) * (sc[k] / sc[0]);
                fv[1] += (mud[i][k] * pp[d2[k + 1]] + mup[i][k] * pd[d2[k + 1]]) * (sc[k] / sc[1]);
                fv[2] += (mud[i][k] * pp[d3[k + 1]] + mup[i][k] * pd[d3[k + 1]]) * (sc[k] / sc[2]);
            }
            for (int d = 0; d < 3; ++d) fv[d] *= sc[d];
            fv[0] += s[i].q * pdp[d1[0]];
            fv[1] += s[i].q * pdp[d2[0]];
            fv[2] += s[i].q * pdp[d3[0]];
            for (int d = 0; d < 3; ++d) F[3 * i + d] -= fv[d] * sc[d] * (0.5 * f);
        }
        energy += 0.5 * f * erind;
        // reciprocal, fixed charges (:2113-2181)
        double erfix = 0.0;
        for (int i = 0; i < n_; ++i) {
            const double* phi = &phi_fixed_[10 * i];
            V[i] += phi[0];
            erfix += s[i].q * phi[0];
            for (int d = 0; d < 3; ++d) F[3 * i + d] -= s[i].q * phi[1 + d] * sc[d] * f;
        }
        energy += 0.5 * f * erfix;
        // self energy (:2472-2508)
        double cii = 0.0;
        for (int i = 0; i < n_; ++i) cii += s[i].q * s[i].q;
        energy += cii * -(f * P_.alpha / (1.0 * el_sqrt_pi));
        for (int i = 0; i < n_; ++i) V[i] += s[i].q * -(P_.alpha / el_sqrt_pi) * 2;
        // charge-derivative chain rule (:2767-2773)
        if (P_.charge_redistribution)
            for (int i = 0; i < n_; ++i)
                for (int q = 0; q < 3; ++q)
                    for (int d = 0; d < 3; ++d) F[3 * i + d] += s[i].dq[q][d] * V[s[i].other[q]] * -f;
        return energy;
    }

   private:
    oracle_elec_params P_;
    int n_;
    double thole_[5];
    V3 box_;
    bool pme_;
    double expgamm_;
    double scale_seconds_ = 0.0, loop_seconds_ = 0.0;  // phase timings of the last run
    int ng_[3] = {0, 0, 0};
    size_t ntot_ = 0;
    std::vector<cplx> grid_;
    std::vector<double> moduli_[3];
    std::vector<double> theta_;
    std::vector<int> igrid_;
    std::vector<double> phi_fixed_, phid_, phip_;
};

}  // namespace

// Full electrostatics for `n` sites; forces accumulated (kJ/mol/nm), returns kJ/mol.
double electrostatics(const oracle_elec_params& P, int n, const double* pos, const double* charge, const int* mol,
                      const int* type, const double* damp, const double* pol, double* forces, oracle_elec_stats* stats,
                      double* dipoles_out, double* fixed_field_out) {
    std::vector<Site> s(n);
    for (int i = 0; i < n; ++i) {
        s[i].pos = load3(pos, i);
        s[i].q = charge[i];
        s[i].mol = mol[i];
        s[i].type = type[i];
        s[i].damp = damp[i];
        s[i].pol = pol[i];
        s[i].other[0] = s[i].other[1] = s[i].other[2] = i;
    }
    Elec e(P, n);
    return e.run(s, forces, stats, dipoles_out, fixed_field_out);
}

// The site charges and converged induced dipoles of the reference's setup()
// (MBPolReferenceElectrostaticsForce.cpp:857-897 and the dipole solve), as
// getElectrostaticPotential and getSystemElectrostaticsMoments compute them
// before their sums; returns false when the dipoles did not converge.
static bool setup_charges_dipoles(const oracle_elec_params& P, int n, const double* pos, const double* charge,
                                  const int* mol, const int* type, const double* damp, const double* pol,
                                  std::vector<double>& q, std::vector<double>& mu) {
    std::vector<Site> s(n);
    for (int i = 0; i < n; ++i) {
        s[i].pos = load3(pos, i);
        s[i].q = charge[i];
        s[i].mol = mol[i];
        s[i].type = type[i];
        s[i].damp = damp[i];
        s[i].pol = pol[i];
        s[i].other[0] = s[i].other[1] = s[i].other[2] = i;
    }
    std::vector<double> forces(3 * (size_t)n, 0.0);
    mu.assign(3 * (size_t)n, 0.0);
    Elec e(P, n);
    try {
        e.run(s, forces.data(), nullptr, mu.data(), nullptr);
    } catch (const std::exception&) {
        return false;
    }
    q.resize(n);
    for (int i = 0; i < n; ++i) q[i] = s[i].q;
    return true;
}

// calculateElectrostaticPotential (:1023-1086): at each grid point, the sum over
// all sites of q/r - mu.(r_site - r_point)/r^3, minimum image with PME
// (getPeriodicDelta :1234-1239, no Ewald split), times the electric constant.
int electrostatic_potential(const oracle_elec_params& P, int n, const double* pos, const double* charge,
                            const int* mol, const int* type, const double* damp, const double* pol, int npts,
                            const double* pts, double* out) {
    std::vector<double> q, mu;
    if (!setup_charges_dipoles(P, n, pos, charge, mol, type, damp, pol, q, mu)) return 1;
    for (int k = 0; k < npts; ++k) out[k] = 0.0;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < npts; ++k) {
            double d[3];
            for (int c = 0; c < 3; ++c) d[c] = pos[3 * i + c] - pts[3 * k + c];
            if (P.method == 1)
                for (int c = 0; c < 3; ++c) d[c] -= std::floor(d[c] * (1.0 / P.box[c]) + 0.5) * P.box[c];
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const double r = std::sqrt(r2);
            const double rr1 = 1.0 / r, rr3 = rr1 * rr1 * rr1;
            out[k] += q[i] * rr1 - (mu[3 * i] * d[0] + mu[3 * i + 1] * d[1] + mu[3 * i + 2] * d[2]) * rr3;
        }
    for (int k = 0; k < npts; ++k) out[k] *= el_electric;
    return 0;
}

// calculateMBPolSystemElectrostaticsMoments (:923-1021): positions relative to
// the centre of mass; out[0] net charge, out[1..3] dipole (Debye), out[4..12]
// the traceless quadrupole (row-major 3x3, Buckingham units).
int system_moments(const oracle_elec_params& P, int n, const double* pos, const double* charge, const int* mol,
                   const int* type, const double* damp, const double* pol, const double* masses, double* out) {
    std::vector<double> q, mu;
    if (!setup_charges_dipoles(P, n, pos, charge, mol, type, damp, pol, q, mu)) return 1;
    double mtot = 0.0, com[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
        mtot += masses[i];
        for (int c = 0; c < 3; ++c) com[c] += pos[3 * i + c] * masses[i];
    }
    if (mtot > 0.0)
        for (int c = 0; c < 3; ++c) com[c] *= 1.0 / mtot;
    double net = 0, dpl[3] = {0, 0, 0}, xx = 0, xy = 0, xz = 0, yy = 0, yz = 0, zz = 0;
    for (int i = 0; i < n; ++i) {
        double r[3];
        for (int c = 0; c < 3; ++c) r[c] = pos[3 * i + c] - com[c];
        const double* m = &mu[3 * i];
        net += q[i];
        for (int c = 0; c < 3; ++c) dpl[c] += r[c] * q[i] + m[c];
        xx += r[0] * r[0] * q[i] + 2.0 * r[0] * m[0];
        xy += r[0] * r[1] * q[i] + r[0] * m[1] + r[1] * m[0];
        xz += r[0] * r[2] * q[i] + r[0] * m[2] + r[2] * m[0];
        yy += r[1] * r[1] * q[i] + 2.0 * r[1] * m[1];
        yz += r[1] * r[2] * q[i] + r[1] * m[2] + r[2] * m[1];
        zz += r[2] * r[2] * q[i] + 2.0 * r[2] * m[2];
    }
    const double qave = (xx + yy + zz) / 3.0;
    out[4] = 0.5 * (xx - qave);
    out[5] = 0.5 * xy;
    out[6] = 0.5 * xz;
    out[8] = 0.5 * (yy - qave);
    out[9] = 0.5 * yz;
    out[12] = 0.5 * (zz - qave);
    out[7] = out[5];
    out[10] = out[6];
    out[11] = out[9];
    const double debye = 4.80321;
    out[0] = net;
    for (int c = 0; c < 3; ++c) out[1 + c] = dpl[c] * 10.0 * debye;
    for (int k = 4; k < 13; ++k) out[k] *= 100.0 * 3.0 * debye;
    return 0;
}

// Thole scale of one pair (white-box known answers).
double thole_scale(const oracle_elec_params& P, double damp_i, double damp_j, int mol_i, int mol_j, int type_i, int type_j,
                   double r, int just_scale, int order, int itype) {
    Elec e(P, 2);
    Site a, b;
    a.damp = damp_i; b.damp = damp_j;
    a.mol = mol_i; b.mol = mol_j;
    a.type = type_i; b.type = type_j;
    return e.scale(a, b, r, just_scale != 0, order, itype);
}

// Charges and charge derivatives of one O,H1,H2,M water (nm).
void water_charges(const double* pos4, double* q4, double* dq) {
    Site s[4];
    for (int i = 0; i < 4; ++i) s[i].pos = load3(pos4, i);
    Elec::water_charge(s[0], s[1], s[2], s[3], 0, 1, 2, 3);
    for (int i = 0; i < 4; ++i) {
        q4[i] = s[i].q;
        for (int q = 0; q < 3; ++q)
            for (int d = 0; d < 3; ++d) dq[(i * 3 + q) * 3 + d] = s[i].dq[q][d];
    }
}

// Direct-sum SOR solve from a given (un-multiplied) fixed field, as exercised by
// the white-box test in TestReferenceMBPolElectrostaticsForce.cpp:136-224.
int solve_given_field(const oracle_elec_params& P, int n, const double* pos, const int* mol, const int* type,
                      const double* damp, const double* pol, const double* field, double* dipoles, oracle_elec_stats* stats) {
    std::vector<Site> s(n);
    std::vector<V3> Ed(n), Ep(n);
    for (int i = 0; i < n; ++i) {
        s[i].pos = load3(pos, i);
        s[i].mol = mol[i];
        s[i].type = type[i];
        s[i].damp = damp[i];
        s[i].pol = pol[i];
        s[i].q = 0;
        Ed[i] = load3(field, i) * pol[i];
        Ep[i] = V3();
    }
    oracle_elec_params Q = P;
    Q.method = 0;
    Elec e(Q, n);
    std::vector<V3> mud = Ed, mup = Ep;
    int it = 0;
    double eps = 0;
    bool conv = e.solve_dipoles(s, Ed, Ep, mud, mup, it, eps);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) dipoles[3 * i + d] = mud[i][d];
    if (stats) { stats->iterations = it; stats->epsilon = eps; stats->converged = conv; }
    return conv ? 0 : 1;
}

}  // namespace oracle
