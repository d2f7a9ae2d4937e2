// 2-body and 3-body permutationally invariant polynomial kernels.
//
// One dimer (trimer) per lane.  Positions are read and minimum-imaged in FP64,
// then made relative to the first oxygen so that the arithmetic type T
// (float on the mixed-precision path, double on the FP64 path) only ever sees
// small coordinates.  The polynomial itself (generated, pip_eval_gen.h) keeps
// variables, gradients and coefficients in registers.  Per-atom forces are
// accumulated in T inside the lane and scattered once as 64-bit fixed point;
// the first molecule's are first summed over the lanes that share it (the
// lists are ordered by first molecule), one atomic per run of lanes.
//
// This file is compiled once per body order and arithmetic type:
//   -DPIP_BODY=2|3  -DPIP_REAL=float|double
// with -DPIP_ORBIT=1 the polynomial is evaluated through the relabelling group
// of the dimer or trimer (pip_orbit_loop.h) instead of term by term.
#include "device_common.h"
// an empty asm that re-defines a variable in its register: no instruction
#define PIP_FENCE(v) asm volatile("" : "+v"(v))
// the same for a uniform pointer in scalar registers
#define PIP_OPAQUE_PTR(p) asm volatile("" : "+s"(p))
// tables read through the constant address space: scalar loads
#define PIP_CONST_PTR(T) const __attribute__((address_space(4))) T*
#if PIP_ORBIT
#include "pip_orbit_loop.h"
#else
#include "pip_eval_gen.h"
#endif
// Register budget of the 3-body kernel, as waves per SIMD (-DPIP3_WAVES=n)
#if defined(PIP3_WAVES) && PIP3_WAVES > 1
#define PIP3_OCCUPANCY __attribute__((amdgpu_waves_per_eu(PIP3_WAVES)))
#else
#define PIP3_OCCUPANCY
#endif

#ifndef PIP_BODY
#error "PIP_BODY must be 2 or 3"
#endif
#ifndef PIP_REAL
#error "PIP_REAL must be float or double"
#endif

namespace mbpol {
namespace {  // kernels have internal linkage: this file is compiled once per body order and type
using namespace mbpol_params;
typedef PIP_REAL real;
typedef v3<real> vr;

__device__ __forceinline__ vr rel(const d3& p, const d3& o) { return vr((real)(p.x - o.x), (real)(p.y - o.y), (real)(p.z - o.z)); }

#if PIP_BODY == 2
// ---------------------------------------------------------------------------
// 2-body: platforms/reference/src/MBPolReferenceTwoBodyForce.cpp:110-296 and
// mbpol_2body_constants.cpp:10-111.  Points (Angstrom, relative to Oa):
// 0 Oa 1 Ha1 2 Ha2 3 Ob 4 Hb1 5 Hb2 6 Xa1 7 Xa2 8 Xb1 9 Xb2.
// ---------------------------------------------------------------------------
struct Var2 {
    int a, b, coul;
    double k, r0;
};
// the 31 variables in reference order (:170-207)
#define V2(a, b, c, k, r) {a, b, c, k, r}
__device__ __forceinline__ Var2 var2(int n) {
    const double di = tb_d0_intra, de = tb_d0_inter;
    switch (n) {
        case 0: return V2(1, 2, 0, tb_k_HH_intra, di);
        case 1: return V2(4, 5, 0, tb_k_HH_intra, di);
        case 2: return V2(0, 1, 0, tb_k_OH_intra, di);
        case 3: return V2(0, 2, 0, tb_k_OH_intra, di);
        case 4: return V2(3, 4, 0, tb_k_OH_intra, di);
        case 5: return V2(3, 5, 0, tb_k_OH_intra, di);
        case 6: return V2(1, 4, 1, tb_k_HH_coul, de);
        case 7: return V2(1, 5, 1, tb_k_HH_coul, de);
        case 8: return V2(2, 4, 1, tb_k_HH_coul, de);
        case 9: return V2(2, 5, 1, tb_k_HH_coul, de);
        case 10: return V2(0, 4, 1, tb_k_OH_coul, de);
        case 11: return V2(0, 5, 1, tb_k_OH_coul, de);
        case 12: return V2(3, 1, 1, tb_k_OH_coul, de);
        case 13: return V2(3, 2, 1, tb_k_OH_coul, de);
        case 14: return V2(0, 3, 1, tb_k_OO_coul, de);
        case 15: return V2(6, 4, 0, tb_k_XH_main, de);
        case 16: return V2(6, 5, 0, tb_k_XH_main, de);
        case 17: return V2(7, 4, 0, tb_k_XH_main, de);
        case 18: return V2(7, 5, 0, tb_k_XH_main, de);
        case 19: return V2(8, 1, 0, tb_k_XH_main, de);
        case 20: return V2(8, 2, 0, tb_k_XH_main, de);
        case 21: return V2(9, 1, 0, tb_k_XH_main, de);
        case 22: return V2(9, 2, 0, tb_k_XH_main, de);
        case 23: return V2(0, 8, 0, tb_k_XO_main, de);
        case 24: return V2(0, 9, 0, tb_k_XO_main, de);
        case 25: return V2(3, 6, 0, tb_k_XO_main, de);
        case 26: return V2(3, 7, 0, tb_k_XO_main, de);
        case 27: return V2(6, 8, 0, tb_k_XX_main, de);
        case 28: return V2(6, 9, 0, tb_k_XX_main, de);
        case 29: return V2(7, 8, 0, tb_k_XX_main, de);
        default: return V2(7, 9, 0, tb_k_XX_main, de);
    }
}
#undef V2

// extra points of one monomer (mbpol_2body_constants.cpp:62-76)
__device__ __forceinline__ void extra_points(const vr& O, const vr& H1, const vr& H2, vr& x1, vr& x2) {
    const vr oh1 = H1 - O, oh2 = H2 - O;
    const vr in_plane = O + (oh1 + oh2) * (real)(0.5 * tb_in_plane_gamma);
    const vr out_of_plane = oh1.cross(oh2) * (real)tb_out_of_plane_gamma;
    x1 = in_plane + out_of_plane;
    x2 = in_plane - out_of_plane;
}

// back-propagation of the extra-point gradients (mbpol_2body_constants.cpp:80-96)
__device__ __forceinline__ void extra_grads(const vr& O, const vr& H1, const vr& H2, const vr& g1, const vr& g2, vr& fO,
                                            vr& fH1, vr& fH2) {
    const vr oh1 = H1 - O, oh2 = H2 - O;
    const vr gm = g1 - g2;
    const vr t1 = oh2.cross(gm), t2 = oh1.cross(gm);
    const vr gsum = g1 + g2;
    const vr in_plane = gsum * (real)(0.5 * tb_in_plane_gamma);
    const vr gh1 = in_plane + t1 * (real)tb_out_of_plane_gamma;
    const vr gh2 = in_plane - t2 * (real)tb_out_of_plane_gamma;
    fO += gsum - (gh1 + gh2);
    fH1 += gh1;
    fH2 += gh2;
}

// Geometry of dimer t: atom indices, imaged positions (Angstrom) and, when the
// O-O distance is inside [2, r2f] Angstrom, the ten points relative to Oa.
// Returns whether the dimer contributes.
__device__ __forceinline__ bool dimer_geometry(long long t, const int2* __restrict__ pairs,
                                               const int* __restrict__ idx, const double* __restrict__ pos,
                                               int periodic, const d3& boxA, int (&at)[6], vr (&R)[10],
                                               d3& dOO, double& rOO) {
    const int2 pr = pairs[t];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        at[q] = idx[3 * pr.x + q];
        at[3 + q] = idx[3 * pr.y + q];
    }
    d3 P[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const d3 p = load_d3(pos, at[q]);
        P[q] = d3{p.x * 10., p.y * 10., p.z * 10.};
    }
    if (periodic) {  // imageMolecules (MBPolReferenceTwoBodyForce.cpp:78-100)
        P[1] = image_to(P[0], P[1], boxA);
        P[2] = image_to(P[0], P[2], boxA);
        P[3] = image_to(P[0], P[3], boxA);
        P[4] = image_to(P[3], P[4], boxA);
        P[5] = image_to(P[3], P[5], boxA);
    }
    dOO = d3{P[0].x - P[3].x, P[0].y - P[3].y, P[0].z - P[3].z};
    rOO = sqrt(dOO.x * dOO.x + dOO.y * dOO.y + dOO.z * dOO.z);
    if (!(rOO <= tb_r2f && rOO >= 2.)) return false;
#pragma unroll
    for (int q = 0; q < 6; ++q) R[q] = rel(P[q], P[0]);
    extra_points(R[0], R[1], R[2], R[6], R[7]);
    extra_points(R[3], R[4], R[5], R[8], R[9]);
    return true;
}

__global__ void __launch_bounds__(64) k_twobody(long long np, const long long* __restrict__ ndev,
                                                const int2* __restrict__ pairs, const int* __restrict__ idx,
                                                const double* __restrict__ pos, int periodic, d3 boxA,
                                                u64* __restrict__ force, u64* __restrict__ energy,
                                                double* __restrict__ eunit) {
    // np: the list's capacity; ndev (when given): its length, known on the device only
    if (ndev) np = min(np, *ndev);
    // a capped grid strides over the list in whole waves
    for (long long t0 = (long long)blockIdx.x * blockDim.x; t0 < np; t0 += (long long)gridDim.x * blockDim.x) {
        const long long t = t0 + threadIdx.x;
        double e_out = 0.0;
        int at[6];
        vr R[10];
        d3 dOO;
        double rOO;
        // forces on the first molecule, summed over the wave's runs of dimers that share it, and
        // on the second, summed over the wave's dimers that share it after a sort by molecule
        u64 fa[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        u64 fb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const long long key = t < np ? (long long)pairs[t].x : -1;
        long long key2 = t < np ? (long long)pairs[t].y : -1;
        if (t < np && dimer_geometry(t, pairs, idx, pos, periodic, boxA, at, R, dOO, rOO)) {
            real x[31], g[31];
#pragma unroll
            for (int n = 0; n < 31; ++n) {
                const Var2 v = var2(n);
                const vr d = R[v.a] - R[v.b];
                const real r = sqrt(d.dot(d));
                const real e = exp_t<real>((real)v.k * ((real)v.r0 - r));
                x[n] = v.coul ? e / r : e;
            }
#if PIP_ORBIT
            real Ep = pip2b_orbit_eval<real>(x, g);
#else
            real Ep = pip2b_eval<real>(x, g);
#endif
            // Only t, Ep and g stay live across the polynomial: the geometry is
            // reloaded from memory (the opaque copy of t keeps the compiler from
            // holding the first loads in registers through the polynomial).
            long long tt = t;
            asm volatile("" : "+v"(tt));
            asm volatile("" : "+v"(Ep));
#pragma unroll
            for (int n = 0; n < 31; ++n) asm volatile("" : "+v"(g[n]));
            dimer_geometry(tt, pairs, idx, pos, periodic, boxA, at, R, dOO, rOO);
            // chain rule through the variables
            vr F[10];
#pragma unroll
            for (int q = 0; q < 10; ++q) F[q] = vr(0, 0, 0);
#pragma unroll
            for (int n = 0; n < 31; ++n) {
                const Var2 v = var2(n);
                const vr d = R[v.a] - R[v.b];
                const real r = sqrt(d.dot(d));
                const real e = exp_t<real>((real)v.k * ((real)v.r0 - r));
                real gg;
                if (v.coul) {
                    const real ri = (real)1 / r, val = e * ri;
                    gg = -((real)v.k + ri) * val * ri;
                } else {
                    gg = -(real)v.k * e / r;
                }
                const vr dd = d * (gg * g[n]);
                F[v.a] += dd;
                F[v.b] -= dd;
            }
            extra_grads(R[0], R[1], R[2], F[6], F[7], F[0], F[1], F[2]);
            extra_grads(R[3], R[4], R[5], F[8], F[9], F[3], F[4], F[5]);
            // switch (mbpol_2body_constants.cpp:98-111), FP64
            double sw, gsw;
            if (rOO > tb_r2i) {
                const double t1 = M_PI / (tb_r2f - tb_r2i);
                const double xs = (rOO - tb_r2i) * t1;
                gsw = -sin(xs) * t1 / 2.0;
                sw = (1.0 + cos(xs)) / 2.0;
            } else {
                sw = 1.0;
                gsw = 0.0;
            }
            const double s = sw * cal2joule * -10.;
            const double gs = gsw * (double)Ep / rOO * cal2joule * -10.;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                double fx = (double)F[q].x * s, fy = (double)F[q].y * s, fz = (double)F[q].z * s;
                if (q == 0) { fx += gs * dOO.x; fy += gs * dOO.y; fz += gs * dOO.z; }
                if (q == 3) { fx -= gs * dOO.x; fy -= gs * dOO.y; fz -= gs * dOO.z; }
                if (q < 3) {
                    fa[3 * q] = to_fix(fx);
                    fa[3 * q + 1] = to_fix(fy);
                    fa[3 * q + 2] = to_fix(fz);
                } else {
                    fb[3 * (q - 3)] = to_fix(fx);
                    fb[3 * (q - 3) + 1] = to_fix(fy);
                    fb[3 * (q - 3) + 2] = to_fix(fz);
                }
            }
            e_out = sw * (double)Ep * cal2joule;
        }
        if (wave_run_sum(key, fa) && key >= 0) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int a = idx[3 * key + q];
#pragma unroll
                for (int d = 0; d < 3; ++d) atomicAdd(&force[3 * a + d], fa[3 * q + d]);
            }
        }
        if (wave_sorted_run_sum(key2, fb) && key2 >= 0) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int a = idx[3 * key2 + q];
#pragma unroll
                for (int d = 0; d < 3; ++d) atomicAdd(&force[3 * a + d], fb[3 * q + d]);
            }
        }
        if (eunit && t < np) eunit[t] = e_out;
        add_energy(energy, e_out);
    }
}

#else
// ---------------------------------------------------------------------------
// 3-body: platforms/reference/src/MBPolReferenceThreeBodyForce.cpp:67-293.
// Points (nm, relative to Oa): 0 Oa 1 Ha1 2 Ha2 3 Ob 4 Hb1 5 Hb2 6 Oc 7 Hc1 8 Hc2.
// ---------------------------------------------------------------------------
struct Var3 {
    int a, b;
    double k, r0;
};
#define V3(a, b, k, r) {a, b, k, r}
__device__ __forceinline__ Var3 var3(int n) {
    switch (n) {  // reference order (:170-206)
        case 0: return V3(1, 2, th_kHH_intra, th_dHH_intra);
        case 1: return V3(4, 5, th_kHH_intra, th_dHH_intra);
        case 2: return V3(7, 8, th_kHH_intra, th_dHH_intra);
        case 3: return V3(0, 1, th_kOH_intra, th_dOH_intra);
        case 4: return V3(0, 2, th_kOH_intra, th_dOH_intra);
        case 5: return V3(3, 4, th_kOH_intra, th_dOH_intra);
        case 6: return V3(3, 5, th_kOH_intra, th_dOH_intra);
        case 7: return V3(6, 7, th_kOH_intra, th_dOH_intra);
        case 8: return V3(6, 8, th_kOH_intra, th_dOH_intra);
        case 9: return V3(1, 4, th_kHH, th_dHH);
        case 10: return V3(1, 5, th_kHH, th_dHH);
        case 11: return V3(1, 7, th_kHH, th_dHH);
        case 12: return V3(1, 8, th_kHH, th_dHH);
        case 13: return V3(2, 4, th_kHH, th_dHH);
        case 14: return V3(2, 5, th_kHH, th_dHH);
        case 15: return V3(2, 7, th_kHH, th_dHH);
        case 16: return V3(2, 8, th_kHH, th_dHH);
        case 17: return V3(4, 7, th_kHH, th_dHH);
        case 18: return V3(4, 8, th_kHH, th_dHH);
        case 19: return V3(5, 7, th_kHH, th_dHH);
        case 20: return V3(5, 8, th_kHH, th_dHH);
        case 21: return V3(0, 4, th_kOH, th_dOH);
        case 22: return V3(0, 5, th_kOH, th_dOH);
        case 23: return V3(0, 7, th_kOH, th_dOH);
        case 24: return V3(0, 8, th_kOH, th_dOH);
        case 25: return V3(3, 1, th_kOH, th_dOH);
        case 26: return V3(3, 2, th_kOH, th_dOH);
        case 27: return V3(3, 7, th_kOH, th_dOH);
        case 28: return V3(3, 8, th_kOH, th_dOH);
        case 29: return V3(6, 1, th_kOH, th_dOH);
        case 30: return V3(6, 2, th_kOH, th_dOH);
        case 31: return V3(6, 4, th_kOH, th_dOH);
        case 32: return V3(6, 5, th_kOH, th_dOH);
        case 33: return V3(0, 3, th_kOO, th_dOO);
        case 34: return V3(0, 6, th_kOO, th_dOO);
        default: return V3(3, 6, th_kOO, th_dOO);
    }
}
#undef V3

__device__ __forceinline__ double switch3(double r, double& g) {  // (:103-120)
    if (r > th_r3f) { g = 0.0; return 0.0; }
    if (r > th_r3i) {
        const double t1 = M_PI / (th_r3f - th_r3i);
        const double x = (r - th_r3i) * t1;
        g = -sin(x) * t1 / 2.0;
        return (1.0 + cos(x)) / 2.0;
    }
    g = 0.0;
    return 1.0;
}

// Geometry of trimer t: atom indices, imaged positions, the three O-O vectors
// (Angstrom) and, unless an O-O distance is below 2 Angstrom (:160-163), the
// nine points in Angstrom relative to Oa.  Returns whether the trimer contributes.
__device__ __forceinline__ bool trimer_geometry(long long t, const int* __restrict__ trip, const int* __restrict__ idx,
                                                const double* __restrict__ pos, int periodic, const d3& box,
                                                int (&at)[9], vr (&R)[9], d3 (&roo)[3], double (&doo)[3]) {
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const int mol = trip[3 * t + m];
#pragma unroll
        for (int q = 0; q < 3; ++q) at[3 * m + q] = idx[3 * mol + q];
    }
    d3 P[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) P[q] = load_d3(pos, at[q]);
    if (periodic) {  // imageMolecules in nm (MBPolReferenceTwoBodyForce.cpp:78-109)
        P[1] = image_to(P[0], P[1], box);
        P[2] = image_to(P[0], P[2], box);
        P[3] = image_to(P[0], P[3], box);
        P[4] = image_to(P[3], P[4], box);
        P[5] = image_to(P[3], P[5], box);
        P[6] = image_to(P[0], P[6], box);
        P[7] = image_to(P[6], P[7], box);
        P[8] = image_to(P[6], P[8], box);
    }
    roo[0] = d3{(P[0].x - P[3].x) * 10., (P[0].y - P[3].y) * 10., (P[0].z - P[3].z) * 10.};  // ab
    roo[1] = d3{(P[0].x - P[6].x) * 10., (P[0].y - P[6].y) * 10., (P[0].z - P[6].z) * 10.};  // ac
    roo[2] = d3{(P[3].x - P[6].x) * 10., (P[3].y - P[6].y) * 10., (P[3].z - P[6].z) * 10.};  // bc
#pragma unroll
    for (int q = 0; q < 3; ++q) doo[q] = sqrt(roo[q].x * roo[q].x + roo[q].y * roo[q].y + roo[q].z * roo[q].z);
    if (doo[0] < 2 || doo[1] < 2 || doo[2] < 2) return false;
    const d3 o{P[0].x * 10., P[0].y * 10., P[0].z * 10.};
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const d3 p{P[q].x * 10., P[q].y * 10., P[q].z * 10.};
        R[q] = rel(p, o);
    }
    return true;
}

__global__ void __launch_bounds__(64) PIP3_OCCUPANCY k_threebody(long long nt, const long long* __restrict__ ndev,
                                                                 const int* __restrict__ trip,
                                                                 const int* __restrict__ idx,
                                                                 const double* __restrict__ pos, int periodic, d3 box,
                                                                 u64* __restrict__ force, u64* __restrict__ energy,
                                                                 double* __restrict__ eunit) {
    // nt: the list's capacity; ndev (when given): its length, known on the device only
    if (ndev) nt = min(nt, *ndev);
    // a capped grid strides over the list in whole waves
    for (long long t0 = (long long)blockIdx.x * blockDim.x; t0 < nt; t0 += (long long)gridDim.x * blockDim.x) {
        const long long t = t0 + threadIdx.x;
        double e_out = 0.0;
        int at[9];
        vr R[9];
        d3 roo[3];
        double doo[3];
        // forces on the first and on the second molecule, each summed over the wave's runs of
        // trimers that share it (the list is ordered by first, then second molecule), and on the
        // third, summed over the wave's trimers that share it after a sort by molecule
        u64 fa[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        u64 fb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        u64 fc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const long long key = t < nt ? (long long)trip[3 * t] : -1;
        const long long key2 = t < nt ? (long long)trip[3 * t + 1] : -1;
        long long key3 = t < nt ? (long long)trip[3 * t + 2] : -1;
        if (t < nt && trimer_geometry(t, trip, idx, pos, periodic, box, at, R, roo, doo)) {
            real x[36], g[36];
#pragma unroll
            for (int n = 0; n < 36; ++n) {
                const Var3 v = var3(n);
                const vr d = R[v.a] - R[v.b];
                const real r = sqrt(d.dot(d));
                x[n] = exp_t<real>(-(real)v.k * (r - (real)v.r0));
            }
#if PIP_ORBIT
            real Ep = pip3b_orbit_eval<real>(x, g);
#else
            real Ep = pip3b_eval<real>(x, g);
#endif
            // Only t, Ep and g stay live across the polynomial; the geometry is
            // reloaded (through an opaque copy of t, so that the first loads are
            // not kept in registers).
            long long tt = t;
            asm volatile("" : "+v"(tt));
            asm volatile("" : "+v"(Ep));
#pragma unroll
            for (int n = 0; n < 36; ++n) asm volatile("" : "+v"(g[n]));
            trimer_geometry(tt, trip, idx, pos, periodic, box, at, R, roo, doo);
            double gab, gac, gbc;
            const double sab = switch3(doo[0], gab), sac = switch3(doo[1], gac), sbc = switch3(doo[2], gbc);
            const double s = sab * sac + sab * sbc + sac * sbc;
            // g_var (:80-101): forces of the variables, times s; units kJ/mol/nm
            vr F[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) F[q] = vr(0, 0, 0);
#pragma unroll
            for (int n = 0; n < 36; ++n) {
                const Var3 v = var3(n);
                const vr d = R[v.a] - R[v.b];
                const real r = sqrt(d.dot(d));
                const real gg = -(real)v.k * g[n] * exp_t<real>(-(real)v.k * (r - (real)v.r0)) / r;
                const vr dd = d * gg;
                F[v.a] += dd;
                F[v.b] -= dd;
            }
            const double fs = s * cal2joule * 10. * -1;
            const double E = (double)Ep;
            gab *= (sac + sbc) * E / doo[0];
            gac *= (sab + sbc) * E / doo[1];
            gbc *= (sab + sac) * E / doo[2];
            const double u = cal2joule * -nm_to_A;
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                double fx = (double)F[q].x * fs, fy = (double)F[q].y * fs, fz = (double)F[q].z * fs;
                if (q == 0) {
                    fx += (gab * roo[0].x + gac * roo[1].x) * u;
                    fy += (gab * roo[0].y + gac * roo[1].y) * u;
                    fz += (gab * roo[0].z + gac * roo[1].z) * u;
                } else if (q == 3) {
                    fx += (gbc * roo[2].x - gab * roo[0].x) * u;
                    fy += (gbc * roo[2].y - gab * roo[0].y) * u;
                    fz += (gbc * roo[2].z - gab * roo[0].z) * u;
                } else if (q == 6) {
                    fx -= (gac * roo[1].x + gbc * roo[2].x) * u;
                    fy -= (gac * roo[1].y + gbc * roo[2].y) * u;
                    fz -= (gac * roo[1].z + gbc * roo[2].z) * u;
                }
                if (q < 3) {
                    fa[3 * q] = to_fix(fx);
                    fa[3 * q + 1] = to_fix(fy);
                    fa[3 * q + 2] = to_fix(fz);
                } else if (q < 6) {
                    fb[3 * (q - 3)] = to_fix(fx);
                    fb[3 * (q - 3) + 1] = to_fix(fy);
                    fb[3 * (q - 3) + 2] = to_fix(fz);
                } else {
                    fc[3 * (q - 6)] = to_fix(fx);
                    fc[3 * (q - 6) + 1] = to_fix(fy);
                    fc[3 * (q - 6) + 2] = to_fix(fz);
                }
            }
            e_out = E * s * cal2joule;
        }
        if (wave_run_sum(key, fa) && key >= 0) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int a = idx[3 * key + q];
#pragma unroll
                for (int d = 0; d < 3; ++d) atomicAdd(&force[3 * a + d], fa[3 * q + d]);
            }
        }
        if (wave_run_sum(key2, fb) && key2 >= 0) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int a = idx[3 * key2 + q];
#pragma unroll
                for (int d = 0; d < 3; ++d) atomicAdd(&force[3 * a + d], fb[3 * q + d]);
            }
        }
        if (wave_sorted_run_sum(key3, fc) && key3 >= 0) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int a = idx[3 * key3 + q];
#pragma unroll
                for (int d = 0; d < 3; ++d) atomicAdd(&force[3 * a + d], fc[3 * q + d]);
            }
        }
        if (eunit && t < nt) eunit[t] = e_out;
        add_energy(energy, e_out);
    }
}
#endif

}  // namespace
}  // namespace mbpol

#define PIP_CAT2(a, b) a##b
#define PIP_CAT(a, b) PIP_CAT2(a, b)
namespace mbpol {
// one 64-lane block per 64 units, at most max_blocks (when positive)
static inline int grid(long long n, int max_blocks) {
    const long long b = (n + 63) / 64;
    return (int)(max_blocks > 0 && b > max_blocks ? max_blocks : b);
}
#if PIP_BODY == 2
void PIP_CAT(launch_twobody_, PIP_SUFFIX)(hipStream_t s, long long npairs, const long long* ndev, const int2* pairs,
                                          const int* idx, const double* pos, int periodic, d3 box, u64* force,
                                          u64* energy, const KEvents& ev, double* eunit, int max_blocks) {
    const d3 boxA{box.x * 10., box.y * 10., box.z * 10.};
    if (npairs > 0)
        MBPOL_LAUNCH(ev, k_twobody, grid(npairs, max_blocks), 64, 0, s, npairs, ndev, pairs, idx, pos, periodic, boxA,
                     force, energy, eunit);
}
#else
void PIP_CAT(launch_threebody_, PIP_SUFFIX)(hipStream_t s, long long ntrip, const long long* ndev, const int* trip,
                                            const int* idx, const double* pos, int periodic, d3 box, u64* force,
                                            u64* energy, const KEvents& ev, double* eunit, int max_blocks) {
    if (ntrip > 0)
        MBPOL_LAUNCH(ev, k_threebody, grid(ntrip, max_blocks), 64, 0, s, ntrip, ndev, trip, idx, pos, periodic, box,
                     force, energy, eunit);
}
#endif
}  // namespace mbpol
