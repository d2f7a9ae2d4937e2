// Per-water and per-particle kernels: virtual-site (M) placement and force
// folding, the per-water terms (1-body surface and geometry-dependent charges,
// water_terms.h), and the fixed-point -> double force/energy finalisation.
// All FP64 (one lane per water; a few thousand flops per water, negligible
// next to the PIPs).
#include <algorithm>

#include "device_common.h"
#include "kernels.h"
#include "water_terms.h"

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

// Per-water terms: the 1-body energy/forces and/or the geometry-dependent
// charges with their derivatives, one water per lane, sharing the geometry
// (water_terms.h).  ONE: 1-body of water w from its O,H1,H2 indices ob[3w..];
// CHG: charges of water w's sites 4w..4w+3 (O, H1, H2, M), q[] written for the
// four sites and dq[w][atom O,H1,H2][site H1,H2,M][xyz] in e/nm.  With both,
// water w of the 1-body list is the water of sites 4w..4w+3 (checked by the
// host), and the two share one geometry unless the 1-body term images its
// hydrogens (periodic) and the imaging moves them.
template <bool ONE, bool CHG>
__global__ void k_water_terms(int nw, const int* __restrict__ ob, const double* __restrict__ pos, int periodic, d3 box,
                              u64* __restrict__ force, u64* __restrict__ energy, double* __restrict__ eunit,
                              double* __restrict__ q, double* __restrict__ dq) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    double e_out = 0.0;
    if (w < nw) {
        const int iO = ONE ? ob[3 * w] : 4 * w, iH1 = ONE ? ob[3 * w + 1] : 4 * w + 1, iH2 = ONE ? ob[3 * w + 2] : 4 * w + 2;
        const d3 pO = load_d3(pos, iO), pH1 = load_d3(pos, iH1), pH2 = load_d3(pos, iH2);
        const double O[3] = {pO.x * 10., pO.y * 10., pO.z * 10.};
        const double H1[3] = {pH1.x * 10., pH1.y * 10., pH1.z * 10.};
        const double H2[3] = {pH2.x * 10., pH2.y * 10., pH2.z * 10.};
        water::Geom g;
        water::geometry(O, H1, H2, g);
        if (CHG) {
            double qw[4], dqw[27];
            water::charges(g, qw, dqw);
#pragma unroll
            for (int k = 0; k < 4; ++k) q[4 * w + k] = qw[k];
            double* out = dq + (size_t)w * 27;
#pragma unroll
            for (int k = 0; k < 27; ++k) out[k] = 10. * dqw[k];  // per nm
        }
        if (ONE) {
            // minimum image of the hydrogens relative to their oxygen
            // (platforms/reference/src/MBPolReferenceOneBodyForce.cpp:204-221)
            double I1[3] = {H1[0], H1[1], H1[2]}, I2[3] = {H2[0], H2[1], H2[2]};
            if (periodic) {
                const d3 q1 = image_to(pO, pH1, box), q2 = image_to(pO, pH2, box);
                I1[0] = q1.x * 10.; I1[1] = q1.y * 10.; I1[2] = q1.z * 10.;
                I2[0] = q2.x * 10.; I2[1] = q2.y * 10.; I2[2] = q2.z * 10.;
                if (I1[0] != H1[0] || I1[1] != H1[1] || I1[2] != H1[2] || I2[0] != H2[0] || I2[1] != H2[1] ||
                    I2[2] != H2[2])
                    water::geometry(O, I1, I2, g);
            }
            const double rHH[3] = {I1[0] - I2[0], I1[1] - I2[1], I1[2] - I2[2]};
            double gH1[3], gH2[3];
            e_out = water::onebody(g, rHH, gH1, gH2);
            // forces in kJ/mol/nm
            add_force(force, iH1, -10. * gH1[0], -10. * gH1[1], -10. * gH1[2]);
            add_force(force, iH2, -10. * gH2[0], -10. * gH2[1], -10. * gH2[2]);
            add_force(force, iO, 10. * (gH1[0] + gH2[0]), 10. * (gH1[1] + gH2[1]), 10. * (gH1[2] + gH2[2]));
            if (eunit) eunit[w] = e_out;
        }
    }
    if (ONE) add_energy(energy, e_out);
}

// Zero `bytes` (a multiple of 4) of device memory with plain stores: a kernel,
// so that it may be captured in a HIP graph.
__global__ void k_fill_zero(unsigned int* __restrict__ p, long long nwords) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nwords;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// zero up to kFillMax buffers in one launch: blockIdx.y picks the buffer
__global__ void k_fill_multi(FillBatch L) {
    const int b = blockIdx.y;
    if (b >= L.n) return;
    unsigned int* p = static_cast<unsigned int*>(L.p[b]);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < L.words[b];
         i += (long long)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// what the host reads after an evaluation, gathered for one copy: the eight
// fixed-point energy sums, then the lengths of the dimer, lower and trimer lists
// (0 for a list not built)
__global__ void k_collect_results(const unsigned long long* __restrict__ energy, const long long* __restrict__ nd,
                                  const int* __restrict__ nl, const long long* __restrict__ nt,
                                  unsigned long long* __restrict__ out) {
    const int t = threadIdx.x;
    if (t < 8) out[t] = energy[t];
    if (t == 8) out[8] = nd ? (unsigned long long)*nd : 0ull;
    if (t == 9) out[9] = nl ? (unsigned long long)(long long)*nl : 0ull;
    if (t == 10) out[10] = nt ? (unsigned long long)*nt : 0ull;
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
                             const double* __restrict__ eunit, u64* __restrict__ out, int comp, long long limit) {
    __shared__ u64 red[256];
    const int g = blockIdx.x;
    // limit: units written (a list's capacity: offsets past it belong to a list that did not fit, and the
    // evaluation is run again)
    const long long b0 = min(limit, off ? off[start[g]] : (long long)mult * start[g]);
    const long long b1 = min(limit, off ? off[start[g + 1]] : (long long)mult * start[g + 1]);
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
                       const double* eunit, u64* out, int comp, long long limit) {
    if (ngroups > 0)
        hipLaunchKernelGGL(k_group_sums, ngroups, 256, 0, s, ngroups, start, mult, off, eunit, out, comp, limit);
}

void launch_water_terms(hipStream_t s, int nw, bool onebody, bool charges, const int* ob, const double* pos,
                        int periodic, d3 box, u64* force, u64* energy, double* eunit, double* q, double* dq) {
    if (nw <= 0 || !(onebody || charges)) return;
    if (onebody && charges)
        hipLaunchKernelGGL((k_water_terms<true, true>), nblk(nw, 64), 64, 0, s, nw, ob, pos, periodic, box, force, energy,
                           eunit, q, dq);
    else if (onebody)
        hipLaunchKernelGGL((k_water_terms<true, false>), nblk(nw, 64), 64, 0, s, nw, ob, pos, periodic, box, force,
                           energy, eunit, q, dq);
    else
        hipLaunchKernelGGL((k_water_terms<false, true>), nblk(nw, 64), 64, 0, s, nw, ob, pos, periodic, box, force,
                           energy, eunit, q, dq);
}
void launch_fill_zero(hipStream_t s, void* p, size_t bytes) {
    const long long nw = (long long)(bytes / 4);
    if (nw > 0)
        hipLaunchKernelGGL(k_fill_zero, (int)std::min<long long>(nblk(nw, 256), 2048), 256, 0, s,
                           static_cast<unsigned int*>(p), nw);
}
void launch_fill_multi(hipStream_t s, const FillBatch& L) {
    long long most = 0;
    for (int b = 0; b < L.n; ++b) most = std::max(most, L.words[b]);
    if (L.n > 0 && most > 0)
        hipLaunchKernelGGL(k_fill_multi, dim3((unsigned)std::min<long long>(nblk(most, 256), 1024), L.n), 256, 0, s, L);
}
void launch_collect_results(hipStream_t s, const u64* energy, const long long* nd, const int* nl, const long long* nt,
                            u64* out) {
    hipLaunchKernelGGL(k_collect_results, 1, 64, 0, s, energy, nd, nl, nt, out);
}
void launch_forces_to_double(hipStream_t s, int n3, const u64* fx, double* f) {
    if (n3 > 0) hipLaunchKernelGGL(k_forces_to_double, nblk(n3, 256), 256, 0, s, n3, fx, f);
}
void launch_fold_vsites(hipStream_t s, int n, const int* vs_idx, const double* vs_w, double* f) {
    if (n > 0) hipLaunchKernelGGL(k_fold_vsites, nblk(n, 256), 256, 0, s, n, vs_idx, vs_w, f);
}
}  // namespace mbpol
