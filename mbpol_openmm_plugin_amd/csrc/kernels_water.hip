// Per-water and per-particle kernels: virtual-site (M) placement and force
// folding, the 1-body monomer surface, the geometry-dependent charges, and the
// fixed-point -> double force/energy finalisation.  All FP64 (one lane per
// water; a few thousand flops per water, negligible next to the PIPs).
#include "device_common.h"
#include "kernels.h"

namespace mbpol {
using namespace mbpol_params;

// ThreeParticleAverageSite placement (python/mbpol.xml:14).
__global__ void k_place_vsites(int n, const int* __restrict__ vs_idx, const double* __restrict__ vs_w,
                               double* __restrict__ pos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = vs_idx[4 * i], a = vs_idx[4 * i + 1], b = vs_idx[4 * i + 2], c = vs_idx[4 * i + 3];
    const double wa = vs_w[3 * i], wb = vs_w[3 * i + 1], wc = vs_w[3 * i + 2];
    for (int d = 0; d < 3; ++d) pos[3 * s + d] = wa * pos[3 * a + d] + wb * pos[3 * b + d] + wc * pos[3 * c + d];
}

// 1-body energy/forces of one water (MBPolReferenceOneBodyForce.cpp:69-221).
__global__ void k_onebody(int nw, const int* __restrict__ idx, const double* __restrict__ pos, int periodic, d3 box,
                          unsigned long long* __restrict__ force, unsigned long long* __restrict__ energy,
                          double* __restrict__ eunit) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    double e_out = 0.0;
    if (w < nw) {
        const int iO = idx[3 * w], iH1 = idx[3 * w + 1], iH2 = idx[3 * w + 2];
        const d3 pO = load_d3(pos, iO);
        d3 pH1 = load_d3(pos, iH1), pH2 = load_d3(pos, iH2);
        if (periodic) {
            pH1 = image_to(pO, pH1, box);
            pH2 = image_to(pO, pH2, box);
        }
        const double r1[3] = {(pH1.x - pO.x) * 10., (pH1.y - pO.y) * 10., (pH1.z - pO.z) * 10.};
        const double r2[3] = {(pH2.x - pO.x) * 10., (pH2.y - pO.y) * 10., (pH2.z - pO.z) * 10.};
        const double rh[3] = {(pH1.x - pH2.x) * 10., (pH1.y - pH2.y) * 10., (pH1.z - pH2.z) * 10.};
        const double d1 = sqrt(r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2]);
        const double d2 = sqrt(r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2]);
        const double dh = sqrt(rh[0] * rh[0] + rh[1] * rh[1] + rh[2] * rh[2]);
        const double costh = (r1[0] * r2[0] + r1[1] * r2[1] + r1[2] * r2[2]) / (d1 * d2);
        const double phh1 = ob_f5z * ob_phh1A * exp(ob_phh2);
        const double e1 = exp(-ob_alphaoh * (d1 - ob_roh));
        const double e2 = exp(-ob_alphaoh * (d2 - ob_roh));
        const double Va = ob_deoh * (e1 * (e1 - 2.0) + e2 * (e2 - 2.0));
        const double Vb = phh1 * exp(-ob_phh2 * dh);
        const double dVa1 = 2.0 * ob_alphaoh * ob_deoh * e1 * (1.0 - e1) / d1;
        const double dVa2 = 2.0 * ob_alphaoh * ob_deoh * e2 * (1.0 - e2) / d2;
        const double dVb = -ob_phh2 * Vb / dh;
        const double x1 = (d1 - ob_reoh) / ob_reoh, x2 = (d2 - ob_reoh) / ob_reoh, x3 = costh - ob_costhe;
        // powers with pw[.][0] = 0 so that (j-1) x^(j-2) vanishes at j = 1
        double pw[3][16];
        pw[0][0] = pw[1][0] = pw[2][0] = 0.0;
        pw[0][1] = pw[1][1] = pw[2][1] = 1.0;
#pragma unroll
        for (int j = 2; j < 16; ++j) {
            pw[0][j] = pw[0][j - 1] * x1;
            pw[1][j] = pw[1][j - 1] * x2;
            pw[2][j] = pw[2][j - 1] * x3;
        }
        const double efac = exp(-ob_b1 * ((d1 - ob_reoh) * (d1 - ob_reoh) + (d2 - ob_reoh) * (d2 - ob_reoh)));
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        for (int j = 1; j < 245; ++j) {
            const int a = ob_idx1[j], b = ob_idx2[j], c = ob_idx3[j];
            const double cz = ob_c5z[j];
            const double pab = pw[0][a] * pw[1][b] + pw[0][b] * pw[1][a];
            s0 += cz * pab * pw[2][c];
            s1 += cz * ((a - 1) * pw[0][a - 1] * pw[1][b] + (b - 1) * pw[0][b - 1] * pw[1][a]) * pw[2][c];
            s2 += cz * ((b - 1) * pw[0][a] * pw[1][b - 1] + (a - 1) * pw[0][b] * pw[1][a - 1]) * pw[2][c];
            s3 += cz * pab * (c - 1) * pw[2][c - 1];
        }
        const double Vc = 2 * ob_c5z[0] + efac * s0;
        double e = Va + Vb + Vc;
        e = (e + ob_e1_shift) * ob_cm1_kcalmol;
        const double dVcdr1 = (-2 * ob_b1 * efac * (d1 - ob_reoh) * s0 + efac * s1 / ob_reoh) / d1;
        const double dVcdr2 = (-2 * ob_b1 * efac * (d2 - ob_reoh) * s0 + efac * s2 / ob_reoh) / d2;
        const double dVcdcth = efac * s3;
        const double u = ob_cm1_kcalmol * cal2joule * 10;
        double g1[3], g2[3];
        for (int i = 0; i < 3; ++i) {
            g1[i] = (dVa1 * r1[i] + dVb * rh[i] + dVcdr1 * r1[i] + dVcdcth * (r2[i] / (d1 * d2) - costh * r1[i] / (d1 * d1))) * u;
            g2[i] = (dVa2 * r2[i] - dVb * rh[i] + dVcdr2 * r2[i] + dVcdcth * (r1[i] / (d1 * d2) - costh * r2[i] / (d2 * d2))) * u;
        }
        add_force(force, iH1, -g1[0], -g1[1], -g1[2]);
        add_force(force, iH2, -g2[0], -g2[1], -g2[2]);
        add_force(force, iO, g1[0] + g2[0], g1[1] + g2[1], g1[2] + g2[2]);
        e_out = e * cal2joule;
        if (eunit) eunit[w] = e_out;
    }
    add_energy(energy, e_out);
}

// Geometry-dependent charges and their derivatives, one O,H1,H2,M water per
// lane (MBPolReferenceElectrostaticsForce.cpp:2793-3096).  q[] is written for
// the four sites; dq[w][atom O,H1,H2][site H1,H2,M][xyz] in e/nm.
__global__ void k_water_charges(int nw, const int* __restrict__ sites, const double* __restrict__ pos,
                                double* __restrict__ q, double* __restrict__ dq) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    // sites == nullptr: water w owns sites 4w..4w+3 (O, H1, H2, M)
    const int iO = sites ? sites[4 * w] : 4 * w, iH1 = sites ? sites[4 * w + 1] : 4 * w + 1;
    const int iH2 = sites ? sites[4 * w + 2] : 4 * w + 2, iM = sites ? sites[4 * w + 3] : 4 * w + 3;
    const d3 O = load_d3(pos, iO), H1 = load_d3(pos, iH1), H2 = load_d3(pos, iH2);
    const double gamma1 = 1.0 - dms_gammaM, gamma2 = dms_gammaM / 2;
    const double xx = dms_Bohr_A, xx2 = xx * xx;
    const double R1[3] = {H1.x * 10. - O.x * 10., H1.y * 10. - O.y * 10., H1.z * 10. - O.z * 10.};
    const double R2[3] = {H2.x * 10. - O.x * 10., H2.y * 10. - O.y * 10., H2.z * 10. - O.z * 10.};
    const double d1 = sqrt(R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2]);
    const double d2 = sqrt(R2[0] * R2[0] + R2[1] * R2[1] + R2[2] * R2[2]);
    const double costh = (R1[0] * R2[0] + R1[1] * R2[1] + R1[2] * R2[2]) / (d1 * d2);
    const double efac = exp(-dms_b1D * ((d1 - dms_reoh) * (d1 - dms_reoh) + (d2 - dms_reoh) * (d2 - dms_reoh)));
    const double x1 = (d1 - dms_reoh) / dms_reoh, x2 = (d2 - dms_reoh) / dms_reoh, x3 = costh - dms_costhe;
    double pw[3][16];
    pw[0][0] = pw[1][0] = pw[2][0] = 0.0;
    pw[0][1] = pw[1][1] = pw[2][1] = 1.0;
#pragma unroll
    for (int j = 2; j < 16; ++j) {
        pw[0][j] = pw[0][j - 1] * x1;
        pw[1][j] = pw[1][j - 1] * x2;
        pw[2][j] = pw[2][j - 1] * x3;
    }
    double p1 = 0, p2 = 0, a11 = 0, a12 = 0, a1c = 0, a21 = 0, a22 = 0, a2c = 0;
    for (int j = 1; j < 84; ++j) {
        const int a = dms_idx0[j], b = dms_idx1[j], c = dms_idx2[j];
        const double cf = dms_coef[j];
        p1 += cf * pw[0][a] * pw[1][b] * pw[2][c];
        p2 += cf * pw[0][b] * pw[1][a] * pw[2][c];
        a11 += cf * (a - 1) * pw[0][a - 1] * pw[1][b] * pw[2][c];
        a12 += cf * (b - 1) * pw[0][a] * pw[1][b - 1] * pw[2][c];
        a1c += cf * (c - 1) * pw[0][a] * pw[1][b] * pw[2][c - 1];
        a21 += cf * (b - 1) * pw[0][b - 1] * pw[1][a] * pw[2][c];
        a22 += cf * (a - 1) * pw[0][b] * pw[1][a - 1] * pw[2][c];
        a2c += cf * (c - 1) * pw[0][b] * pw[1][a] * pw[2][c - 1];
    }
    a11 /= dms_reoh / xx;
    a12 /= dms_reoh / xx;
    a21 /= dms_reoh / xx;
    a22 /= dms_reoh / xx;
    const double pl1 = costh, pl2 = 0.5 * (3 * pl1 * pl1 - 1.0);
    const double ang = dms_c0 + pl1 * dms_c1 + pl2 * dms_c2;
    const double pc0 = dms_a * (pow(d1, dms_b) + pow(d2, dms_b)) * ang;
    const double dpc0dr1 = dms_a * dms_b * pow(d1, dms_b - 1) * ang * xx2;
    const double dpc0dr2 = dms_a * dms_b * pow(d2, dms_b - 1) * ang * xx2;
    const double dpc0dc = dms_a * (pow(d1, dms_b) + pow(d2, dms_b)) * (dms_c1 + 0.5 * (6.0 * pl1) * dms_c2) * xx;
    const double defdr1 = -2.0 * dms_b1D * (d1 - dms_reoh) * efac * xx;
    const double defdr2 = -2.0 * dms_b1D * (d2 - dms_reoh) * efac * xx;
    a11 = a11 * efac + p1 * defdr1 + dpc0dr1;
    a12 = a12 * efac + p1 * defdr2 + dpc0dr2;
    a1c = a1c * efac + dpc0dc;
    a21 = a21 * efac + p2 * defdr1 + dpc0dr1;
    a22 = a22 * efac + p2 * defdr2 + dpc0dr2;
    a2c = a2c * efac + dpc0dc;
    p1 = dms_coef[0] + p1 * efac + pc0 * xx;
    p2 = dms_coef[0] + p2 * efac + pc0 * xx;
    const double g21 = gamma2 / gamma1;
    q[iO] = 0.;
    q[iH1] = p1 + g21 * (p1 + p2);
    q[iH2] = p2 + g21 * (p1 + p2);
    q[iM] = -(p1 + p2) / gamma1;
    a11 /= xx;
    a12 /= xx;
    a21 /= xx;
    a22 /= xx;
    const double f1q1r13 = (a11 - (a1c * costh / d1)) / d1, f1q1r23 = a1c / (d1 * d2);
    const double f2q1r23 = (a12 - (a1c * costh / d2)) / d2, f2q1r13 = a1c / (d2 * d1);
    const double f1q2r13 = (a21 - (a2c * costh / d1)) / d1, f1q2r23 = a2c / (d1 * d2);
    const double f2q2r23 = (a22 - (a2c * costh / d2)) / d2, f2q2r13 = a2c / (d2 * d1);
    double* out = dq + (size_t)w * 27;  // [atom O,H1,H2][site H1,H2,M][xyz]
    for (int i = 0; i < 3; ++i) {
        const double q1H1 = f1q1r13 * R1[i] + f1q1r23 * R2[i], q1H2 = f2q1r13 * R1[i] + f2q1r23 * R2[i];
        const double q1O = -(q1H1 + q1H2);
        const double q2H1 = f1q2r13 * R1[i] + f1q2r23 * R2[i], q2H2 = f2q2r13 * R1[i] + f2q2r23 * R2[i];
        const double q2O = -(q2H1 + q2H2);
        const double sH1 = g21 * (q1H1 + q2H1), sH2 = g21 * (q1H2 + q2H2), sO = g21 * (q1O + q2O);
        out[(0 * 3 + 0) * 3 + i] = 10 * (q1O + sO);
        out[(1 * 3 + 0) * 3 + i] = 10 * (q1H1 + sH1);
        out[(2 * 3 + 0) * 3 + i] = 10 * (q1H2 + sH2);
        out[(0 * 3 + 1) * 3 + i] = 10 * (q2O + sO);
        out[(1 * 3 + 1) * 3 + i] = 10 * (q2H1 + sH1);
        out[(2 * 3 + 1) * 3 + i] = 10 * (q2H2 + sH2);
        out[(0 * 3 + 2) * 3 + i] = 10 * (-(q1O + q2O) - 2 * sO);
        out[(1 * 3 + 2) * 3 + i] = 10 * (-(q1H1 + q2H1) - 2 * sH1);
        out[(2 * 3 + 2) * 3 + i] = 10 * (-(q1H2 + q2H2) - 2 * sH2);
    }
}

// fixed point -> double for every particle
__global__ void k_forces_to_double(int n3, const unsigned long long* __restrict__ fx, double* __restrict__ f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) f[i] = (double)(long long)fx[i] * kFixInv;
}

// Fold each virtual site's force onto its parents with the placement weights and zero it.
__global__ void k_fold_vsites(int n, const int* __restrict__ vs_idx, const double* __restrict__ vs_w,
                              double* __restrict__ f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = vs_idx[4 * i];
    const double fs[3] = {f[3 * s], f[3 * s + 1], f[3 * s + 2]};
    for (int k = 0; k < 3; ++k) {
        const int p = vs_idx[4 * i + 1 + k];
        const double wk = vs_w[3 * i + k];
        for (int d = 0; d < 3; ++d) atomicAdd(&f[3 * p + d], wk * fs[d]);
    }
    f[3 * s] = f[3 * s + 1] = f[3 * s + 2] = 0.0;
}

}  // namespace mbpol

namespace mbpol {
static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }
void launch_place_vsites(hipStream_t s, int n, const int* vs_idx, const double* vs_w, double* pos) {
    if (n > 0) hipLaunchKernelGGL(k_place_vsites, nblk(n, 256), 256, 0, s, n, vs_idx, vs_w, pos);
}
__global__ void k_group_sums(int ngroups, const int* __restrict__ start, int mult, const long long* __restrict__ off,
                             const double* __restrict__ eunit, u64* __restrict__ out, int comp) {
    __shared__ u64 red[256];
    const int g = blockIdx.x;
    const long long b0 = off ? off[start[g]] : (long long)mult * start[g];
    const long long b1 = off ? off[start[g + 1]] : (long long)mult * start[g + 1];
    u64 s = 0;
    for (long long u = b0 + threadIdx.x; u < b1; u += 256) s += to_fix(eunit[u]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)g * 8 + comp] = red[0];
}

void launch_group_sums(hipStream_t s, int ngroups, const int* start, int mult, const long long* off,
                       const double* eunit, u64* out, int comp) {
    if (ngroups > 0) hipLaunchKernelGGL(k_group_sums, ngroups, 256, 0, s, ngroups, start, mult, off, eunit, out, comp);
}

void launch_onebody(hipStream_t s, int nw, const int* idx, const double* pos, int periodic, d3 box, u64* force,
                    u64* energy, double* eunit) {
    if (nw > 0)
        hipLaunchKernelGGL(k_onebody, nblk(nw, 64), 64, 0, s, nw, idx, pos, periodic, box, force, energy, eunit);
}
void launch_water_charges(hipStream_t s, int nw, const int* sites, const double* pos, double* q, double* dq) {
    if (nw > 0) hipLaunchKernelGGL(k_water_charges, nblk(nw, 64), 64, 0, s, nw, sites, pos, q, dq);
}
void launch_forces_to_double(hipStream_t s, int n3, const u64* fx, double* f) {
    if (n3 > 0) hipLaunchKernelGGL(k_forces_to_double, nblk(n3, 256), 256, 0, s, n3, fx, f);
}
void launch_fold_vsites(hipStream_t s, int n, const int* vs_idx, const double* vs_w, double* f) {
    if (n > 0) hipLaunchKernelGGL(k_fold_vsites, nblk(n, 256), 256, 0, s, n, vs_idx, vs_w, f);
}
}  // namespace mbpol
