// Thole-damped polarisable electrostatics: real-space pair kernels, the PME
// charge/dipole spreading and potential gathering, the SOR dipole update and
// the energy/force kernels.
//
// Restates platforms/reference/src/MBPolReferenceElectrostaticsForce.cpp.  The
// reference carries two dipole sets ("d" and "p") that start equal and are
// updated identically (:618-647, :516-531); here one set is kept and every
// d/p-symmetric expression is written for d = p.
//
// Pair kernels are "site-centric": lanes of one group of G lanes share a site
// i and stride over its neighbour rows, so every contribution to site i stays
// in registers and is reduced across the group.  Neighbour rows, bins and
// spreading are ordered by site index, so results do not depend on scheduling.
// Sites of zero polarisability keep zero induced dipoles throughout; the
// iteration therefore runs over the polarisable sites only.
#include <algorithm>

#include "device_common.h"
#include "elec_kernels.h"

namespace mbpol {
using namespace mbpol_params;

#pragma clang fp contract(fast)

// ---------------------------------------------------------------------------
// Thole damping (getAndScaleInverseRs, :260-328)
// ---------------------------------------------------------------------------
// regularised upper incomplete gamma Q(3/4, x): series below a+1, Lentz
// continued fraction above (Press et al., Numerical Recipes 3rd ed. 6.2)
__device__ __noinline__ double gammq34_eval(double x) {
    const double a = 0.75;
    const double eps = 2.220446049250313e-16, fpmin = 2.2250738585072014e-308 / 2.220446049250313e-16;
    const double lg = 0.20328095143129538;  // ln Gamma(3/4) (Gamma(3/4) = 1.2254...)
    if (x == 0.0) return 1.0;
    if (x > 745.0) return 0.0;  // exp(-x) underflows: the result is 0 to double precision
    if (x < a + 1.0) {
        double ap = a, sum = 1.0 / a, del = sum;
        for (int it = 0; it < 200; ++it) {
            ap += 1.0;
            del *= x / ap;
            sum += del;
            if (fabs(del) < fabs(sum) * eps) break;
        }
        return 1.0 - sum * exp(-x + a * log(x) - lg);
    }
    double b = x + 1.0 - a, c = 1.0 / fpmin, d = 1.0 / b, h = d;
    for (int i = 1; i < 400; ++i) {
        const double an = -i * (i - a);
        b += 2.0;
        d = an * d + b;
        if (fabs(d) < fpmin) d = fpmin;
        c = b + an / c;
        if (fabs(c) < fpmin) c = fpmin;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (fabs(del - 1.0) <= eps) break;
    }
    return exp(-x + a * log(x) - lg) * h;
}
// out of line: most pairs of a cutoff sphere have x > 745, where Q is 0
__device__ __forceinline__ double gammq34(double x) { return x > 745.0 ? 0.0 : gammq34_eval(x); }

constexpr double kGamma34 = 1.2254167024651776451;  // Gamma(3/4)

// Thole scale factors of one pair (justScale = true forms).  u = (r/A)^4 with
// A = (damp_i damp_j)^(1/6), formed as the product of the per-site sixth roots
// a_i = damp_i^(1/6); A == 0 means undamped (all scales 1).  The scales of one
// Thole parameter g share e = exp(-g u): ex(g) forms it once per pair.
struct Thole {
    double A, ur, ratio;  // ur = r / A, ratio = ur^4
    __device__ __forceinline__ Thole(double ai, double aj, double r) {
        A = ai * aj;
        ur = (A != 0.0) ? r / A : 0.0;
        const double u2 = ur * ur;
        ratio = u2 * u2;
    }
    __device__ __forceinline__ bool on() const { return A != 0.0 && A > -50.0; }
    __device__ __forceinline__ double ex(double g) const { return exp(-g * ratio); }
    // gq = g^(1/4)
    __device__ __forceinline__ double s1(double g, double gq, double e) const {
        if (!on()) return 1.0;
        return 1.0 - e + gq * ur * kGamma34 * gammq34(g * ratio);
    }
    __device__ __forceinline__ double s3(double e) const { return on() ? 1.0 - e : 1.0; }
    __device__ __forceinline__ double s5(double g, double e) const {
        if (!on()) return 1.0;
        return (1.0 - e) - (4. / 3.) * g * e * ratio;
    }
    // (r/A)^4 stands for the reference's r^4 / A^4
    __device__ __forceinline__ double s7(double g, double e) const {
        if (!on()) return 1.0;
        return ((1.0 - e) - (4. / 3.) * g * e * ratio) - (4. / 15.) * g * (4. * g * ratio - 1.) * e * ratio;
    }
};

// Thole parameter for the dipole-dipole interaction (:289-307)
__device__ __forceinline__ double tdd_gamma(const ElecParams& P, int same, int ti, int tj) {
    if (same) return (ti == 0 || tj == 0) ? P.thole[3] : P.thole[4];
    return P.thole[2];
}

// Ewald real-space screening functions bn0..bn3 (:2541-2557); alpha > 0.  The
// reference's divisions by r and r^2 are multiplications by ri = 1/r and ri^2
// (one division per pair, shared with the caller's own 1/r).
struct Ewald {
    double bn0, bn1, bn2, bn3;
    __device__ __forceinline__ Ewald(double alpha, double r) {
        const double ra = alpha * r;
        const double ri = 1.0 / r, r2i = ri * ri;
        bn0 = erfc(ra) * ri;
        const double alsq2 = 2.0 * alpha * alpha;
        double alsq2n = 1.0 / (el_sqrt_pi * alpha);
        const double ex = exp(-(ra * ra));
        alsq2n *= alsq2;
        bn1 = (bn0 + alsq2n * ex) * r2i;
        alsq2n *= alsq2;
        bn2 = (3.0 * bn1 + alsq2n * ex) * r2i;
        alsq2n *= alsq2;
        bn3 = (5.0 * bn2 + alsq2n * ex) * r2i;
    }
};

__device__ __forceinline__ void min_image(double& dx, double& dy, double& dz, const ElecParams& P) {
    if (!P.pme) return;
    dx -= floor(dx * P.inv_box.x + 0.5) * P.box.x;
    dy -= floor(dy * P.inv_box.y + 0.5) * P.box.y;
    dz -= floor(dz * P.inv_box.z + 0.5) * P.box.z;
}



template <int G>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
    return v;
}

// bits of the G-lane group that contains this lane, and of the lanes below it
template <int G>
__device__ __forceinline__ unsigned long long group_bits() {
    const int lane = threadIdx.x & 63;
    return G == 64 ? ~0ull : (((1ull << G) - 1ull) << (lane & ~(G - 1)));
}
__device__ __forceinline__ unsigned long long lanes_below() { return (1ull << (threadIdx.x & 63)) - 1ull; }

// ---------------------------------------------------------------------------
// site set-up and neighbour rows
// ---------------------------------------------------------------------------
__global__ void k_site_charges_copy(int n, const double* __restrict__ q0, double* __restrict__ q) {
    critical_path_wave();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = q0[i];
}

// a copy that happens only when the solve stopped converged (the fixed-length
// solve of a captured evaluation keeps the previous dipoles otherwise)
__global__ void k_copy_if_converged(long long n, const SolveCtl* __restrict__ ctl, const double* __restrict__ src,
                                    double* __restrict__ dst) {
    if (!(ctl->stop && ctl->converged)) return;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_site_cells(int n, const double* __restrict__ pos, ElecParams P, int3 nc, int* __restrict__ cell_of,
                             int* __restrict__ cell_count) {
    critical_path_wave();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    const double q[3] = {x - P.box.x * floor(x * P.inv_box.x), y - P.box.y * floor(y * P.inv_box.y),
                         z - P.box.z * floor(z * P.inv_box.z)};
    const double inv[3] = {P.inv_box.x, P.inv_box.y, P.inv_box.z};
    const int ncd[3] = {nc.x, nc.y, nc.z};
    int c[3];
    for (int d = 0; d < 3; ++d) {
        int v = (int)floor(q[d] * inv[d] * ncd[d]);
        c[d] = v < 0 ? 0 : (v >= ncd[d] ? ncd[d] - 1 : v);
    }
    const int cell = (c[0] * nc.y + c[1]) * nc.z + c[2];
    cell_of[i] = cell;
    atomicAdd(&cell_count[cell], 1);
}

// scatter item i into its bucket (slot order is arbitrary; k_sort_segments restores index order)
__global__ void k_bucket_fill(int n, const int* __restrict__ bucket_of, const int* __restrict__ bucket_start,
                              int* __restrict__ bucket_fill, int* __restrict__ members) {
    critical_path_wave();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = bucket_of[i];
    if (c < 0) return;  // item without a bucket
    members[bucket_start[c] + atomicAdd(&bucket_fill[c], 1)] = i;
}

// Members of every bucket in ascending item order, one lane per item: the
// item's slot is the number of smaller items in its bucket (members_in holds
// each bucket's items in arbitrary order).  The order of every neighbour row
// and spreading sum then depends on the input only.
__global__ void k_rank_sort(int n, const int* __restrict__ bucket_of, const int* __restrict__ bucket_start,
                            const int* __restrict__ members_in, int* __restrict__ members_out) {
    critical_path_wave();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = bucket_of[i];
    if (c < 0) return;
    const int b = bucket_start[c], e = bucket_start[c + 1];
    int rank = 0;
    for (int k = b; k < e; ++k) rank += members_in[k] < i;
    members_out[b + rank] = i;
}

// Neighbour rows of every site, in fixed-capacity (ELL) layout: row i of the
// P table lists the polarisable sites j != i with r_ij <= rc (minimum image
// under PME), row i of the Q table the non-polarisable ones.  One group of G
// lanes per site; candidates are taken in cell order and compacted with a
// wavefront ballot.  Counts are kept beyond the capacity so that the host can
// detect an overflow (maxcnt[0..1]) and rebuild with larger rows; maxcnt[2]
// accumulates the number of directed pairs, maxcnt[3] the P entries of the
// polarisable sites (the rows the dipole iteration streams).
// positions and polarisable flags in cell order: the row build reads its
// candidates from consecutive entries instead of gathering them by site index
__global__ void k_site_pack(int n, const int* __restrict__ members, const double* __restrict__ pos,
                            const double* __restrict__ pol, double4* __restrict__ cpos) {
    critical_path_wave();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int j = members[t];
    cpos[t] = make_double4(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], pol[j] != 0.0 ? 1.0 : 0.0);
}

template <int G>
__global__ void __launch_bounds__(256) k_site_lists(int n, const double* __restrict__ pos, const double* __restrict__ pol,
                                                    ElecParams P, int3 nc, int reach, int use_cells,
                                                    const int* __restrict__ cell_of, const int* __restrict__ cell_start,
                                                    const int* __restrict__ members,
                                                    const double4* __restrict__ cpos, int capP, int capQ,
                                                    int* __restrict__ nbP, int* __restrict__ cntP,
                                                    int* __restrict__ nbQ, int* __restrict__ cntQ,
                                                    int* __restrict__ maxcnt) {
    critical_path_wave();
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / G, lane = threadIdx.x % G;
    int cP = 0, cQ = 0;
    if (i < n) {  // uniform over the group
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const double rc2 = P.cutoff * P.cutoff;
    const unsigned long long gb = group_bits<G>(), below = lanes_below();
    int* rowP = nbP + (size_t)i * capP;
    int* rowQ = nbQ + (size_t)i * capQ;
    // q: position of j and 1 if j is polarisable (from the cell-ordered copy when cells are used)
    auto consider = [&](int j, bool active, const double4& q) {
        bool ok = false, pj = false;
        if (active && j != i) {
            double dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
            min_image(dx, dy, dz, P);
            ok = dx * dx + dy * dy + dz * dz <= rc2;
            pj = q.w != 0.0;
        }
        const bool inP = ok && pj, inQ = ok && !pj;
        const unsigned long long bP = __ballot(inP) & gb, bQ = __ballot(inQ) & gb;
        if (inP) {
            const int k = cP + __popcll(bP & below);
            if (k < capP) rowP[k] = j;
        }
        if (inQ) {
            const int k = cQ + __popcll(bQ & below);
            if (k < capQ) rowQ[k] = j;
        }
        cP += __popcll(bP);
        cQ += __popcll(bQ);
    };
    if (!use_cells) {
        const int b0 = P.gs ? P.gs[i] : 0, b1 = P.gs ? P.ge[i] : n;
        for (int b = b0; b < b1; b += G) {
            const int j = b + lane;
            const bool act = j < b1;
            const double4 q = act ? make_double4(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], pol[j] != 0.0 ? 1.0 : 0.0)
                                  : make_double4(0.0, 0.0, 0.0, 0.0);
            consider(j, act, q);
        }
    } else {
        // cells (X, Y, Z - reach .. Z + reach) are consecutive in the cell order, so each
        // (X, Y) column is one or two contiguous member ranges, visited in the same order
        // as cell by cell
        const int cell = cell_of[i];
        const int cz = cell % nc.z, cy = (cell / nc.z) % nc.y, cx = cell / (nc.z * nc.y);
        auto scan = [&](int base, int za, int zb) {
            const int b0 = cell_start[base + za], b1 = cell_start[base + zb + 1];
            for (int t = b0; t < b1; t += G) {
                const int tt = t + lane;
                const bool act = tt < b1;
                consider(act ? members[tt] : -1, act, act ? cpos[tt] : make_double4(0.0, 0.0, 0.0, 0.0));
            }
        };
        for (int ax = -reach; ax <= reach; ++ax)
            for (int ay = -reach; ay <= reach; ++ay) {
                int X = cx + ax, Y = cy + ay;
                X = X < 0 ? X + nc.x : (X >= nc.x ? X - nc.x : X);
                Y = Y < 0 ? Y + nc.y : (Y >= nc.y ? Y - nc.y : Y);
                const int base = (X * nc.y + Y) * nc.z, z0 = cz - reach, z1 = cz + reach;
                if (z0 < 0) {
                    scan(base, z0 + nc.z, nc.z - 1);
                    scan(base, 0, z1);
                } else if (z1 >= nc.z) {
                    scan(base, z0, nc.z - 1);
                    scan(base, 0, z1 - nc.z);
                } else {
                    scan(base, z0, z1);
                }
            }
    }
    if (lane == 0) {
        cntP[i] = cP;
        cntQ[i] = cQ;
    }
    }
    // row maxima and totals: combined in LDS, one set of global atomics per block
    __shared__ int red[4];
    if (threadIdx.x < 4) red[threadIdx.x] = 0;
    __syncthreads();
    if (i < n && lane == 0) {
        atomicMax(&red[0], cP);
        atomicMax(&red[1], cQ);
        atomicAdd(&red[2], cP + cQ);
        if (pol[i] != 0.0) atomicAdd(&red[3], cP);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicMax(&maxcnt[0], red[0]);
        atomicMax(&maxcnt[1], red[1]);
        atomicAdd(&maxcnt[2], red[2]);
        atomicAdd(&maxcnt[3], red[3]);
    }
}

// visit every neighbour of site i (P row, then Q row) with G lanes
template <int G, typename F>
__device__ __forceinline__ void for_rows(int i, int lane, int capP, const int* __restrict__ nbP,
                                         const int* __restrict__ cntP, int capQ, const int* __restrict__ nbQ,
                                         const int* __restrict__ cntQ, F&& f) {
    const int cp = min(cntP[i], capP), cq = min(cntQ[i], capQ);
    const int* rp = nbP + (size_t)i * capP;
    const int* rq = nbQ + (size_t)i * capQ;
    for (int k = lane; k < cp; k += G) f(rp[k]);
    for (int k = lane; k < cq; k += G) f(rq[k]);
}

// One pass over the rows of polarisable site i:
// - induced-field prefactors of every (polarisable i, polarisable j) entry of
//   the P row, field_i += p1 mu_j + p2 (mu_j.d) d (PME :2374-2390 with the
//   precomputed justScale factors of :2280-2298; direct sum :438-486 with
//   :534-555), stored per entry;
// - the real-space fixed field at i over the P and Q rows (:361-420 direct
//   sum, :1342-1407 PME), summed per lane in row order and added to field[i].
// The P-row pairs share the distance, the Ewald factors and the Thole sixth root.
template <int G, bool PME>
__global__ void __launch_bounds__(256) k_prefactors_field(int npol, const int* __restrict__ polidx,
                                                          const double* __restrict__ pos, const SiteParams S,
                                                          ElecParams P, int capP, const int* __restrict__ nbP,
                                                          const int* __restrict__ cntP, int capQ,
                                                          const int* __restrict__ nbQ, const int* __restrict__ cntQ,
                                                          double2* __restrict__ pf, double* __restrict__ field) {
    critical_path_wave();
    const int g = (blockIdx.x * blockDim.x + threadIdx.x) / G, lane = threadIdx.x % G;
    if (g >= npol) return;
    const int i = polidx[g];
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    double ex = 0, ey = 0, ez = 0;
    // fixed-field term of pair (i, j) at distance r
    auto field_term = [&](int j, double dx, double dy, double dz, double r2, double r, const Thole& th, bool same,
                          double bn1) {
        if constexpr (!PME) {
            if (same) return;
            const double ri = 1.0 / r;
            const double rr3 = th.s3(th.ex(P.thole[0])) * (ri * ri * ri);
            const double f = -rr3 * S.q[j];
            ex += dx * f;
            ey += dy * f;
            ez += dz * f;
        } else {
            double s3 = th.s3(th.ex(P.thole[0]));
            if (same) s3 = 2;  // reproduced as in the reference (:1382-1385)
            const double ri = 1.0 / r;
            const double rr3 = (s3 - 1.) * (ri * ri * ri);
            const double f = -(bn1 - rr3) * S.q[j];
            ex += dx * f;
            ey += dy * f;
            ez += dz * f;
        }
    };
    {
        const int c = min(cntP[i], capP);
        const int* row = nbP + (size_t)i * capP;
        double2* prow = pf + (size_t)i * capP;
        for (int k = lane; k < c; k += G) {
            const int j = row[k];
            double dx = pos[3 * j] - xi, dy = pos[3 * j + 1] - yi, dz = pos[3 * j + 2] - zi;
            if constexpr (PME) min_image(dx, dy, dz, P);
            const double r2 = dx * dx + dy * dy + dz * dz, r = sqrt(r2);
            const Thole th(S.damp6[i], S.damp6[j], r);
            const bool same = S.mol[i] == S.mol[j];
            const double gam = tdd_gamma(P, same, S.type[i], S.type[j]);
            const double eg = th.ex(gam);
            const double sc3 = th.s3(eg), sc5 = th.s5(gam, eg);
            double p1, p2, bn1 = 0.0;
            if constexpr (PME) {
                const Ewald ew(P.alpha, r);
                const double ri = 1.0 / r, ri3 = ri * ri * ri;
                p1 = (1.0 - sc3) * ri3 - ew.bn1;
                p2 = ew.bn2 - 3.0 * (1.0 - sc5) * (ri3 * ri * ri);
                bn1 = ew.bn1;
            } else {
                const double ri = 1.0 / r, r2i = ri * ri;
                p1 = -(ri * r2i) * sc3;
                p2 = (ri * r2i * 3.0 * r2i) * sc5;
            }
            prow[k] = make_double2(p1, p2);
            field_term(j, dx, dy, dz, r2, r, th, same, bn1);
        }
    }
    {
        const int c = min(cntQ[i], capQ);
        const int* row = nbQ + (size_t)i * capQ;
        for (int k = lane; k < c; k += G) {
            const int j = row[k];
            double dx = pos[3 * j] - xi, dy = pos[3 * j + 1] - yi, dz = pos[3 * j + 2] - zi;
            if constexpr (PME) min_image(dx, dy, dz, P);
            const double r2 = dx * dx + dy * dy + dz * dz, r = sqrt(r2);
            const Thole th(S.damp6[i], S.damp6[j], r);
            double bn1 = 0.0;
            if constexpr (PME) bn1 = Ewald(P.alpha, r).bn1;
            field_term(j, dx, dy, dz, r2, r, th, S.mol[i] == S.mol[j], bn1);
        }
    }
    ex = group_sum<G>(ex);
    ey = group_sum<G>(ey);
    ez = group_sum<G>(ez);
    if (lane == 0) {
        field[3 * i] += ex;
        field[3 * i + 1] += ey;
        field[3 * i + 2] += ez;
    }
}

// ---------------------------------------------------------------------------
// PME (:1446-2111)
// ---------------------------------------------------------------------------
// order-5 B-spline values and derivatives (computeBSplinePoint, :1446-1519)
__device__ __forceinline__ void bspline5(double w, double th[5][4]) {
    double A[6][6];
    A[2][2] = w;
    A[2][1] = 1.0 - w;
    A[3][3] = 0.5 * w * A[2][2];
    A[3][2] = 0.5 * ((1.0 + w) * A[2][1] + (2.0 - w) * A[2][2]);
    A[3][1] = 0.5 * (1.0 - w) * A[2][1];
#pragma unroll
    for (int i = 4; i <= 5; ++i) {
        const int k = i - 1;
        const double den = 1.0 / k;
        A[i][i] = den * w * A[k][k];
#pragma unroll
        for (int j = 1; j <= i - 2; ++j) A[i][i - j] = den * ((w + j) * A[k][i - j - 1] + (i - j - w) * A[k][i - j]);
        A[i][1] = den * (1.0 - w) * A[k][1];
    }
    // derivative coefficient rows (differences of the lower-order rows)
    A[4][5] = A[4][4];
#pragma unroll
    for (int i = 4; i >= 2; --i) A[4][i] = A[4][i - 1] - A[4][i];
    A[4][1] = -A[4][1];
    A[3][4] = A[3][3];
#pragma unroll
    for (int i = 3; i >= 2; --i) A[3][i] = A[3][i - 1] - A[3][i];
    A[3][1] = -A[3][1];
    A[3][5] = A[3][4];
#pragma unroll
    for (int i = 4; i >= 2; --i) A[3][i] = A[3][i - 1] - A[3][i];
    A[3][1] = -A[3][1];
    A[2][3] = A[2][2];
    A[2][2] = A[2][1] - A[2][2];
    A[2][1] = -A[2][1];
    A[2][4] = A[2][3];
#pragma unroll
    for (int i = 3; i >= 2; --i) A[2][i] = A[2][i - 1] - A[2][i];
    A[2][1] = -A[2][1];
    A[2][5] = A[2][4];
#pragma unroll
    for (int i = 4; i >= 2; --i) A[2][i] = A[2][i - 1] - A[2][i];
    A[2][1] = -A[2][1];
#pragma unroll
    for (int i = 1; i <= 5; ++i) {
        th[i - 1][0] = A[5][i];
        th[i - 1][1] = A[4][i];
        th[i - 1][2] = A[3][i];
        th[i - 1][3] = A[2][i];
    }
}


// grid origin g0 and fractional offsets of one site along each axis
// (computeMBPolBsplines, :1524-1554); the single place both are computed
__device__ __forceinline__ void site_origin(const double* pos, int i, const ElecParams& P, int g0[3], double w[3]) {
    double p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    const double L[3] = {P.box.x, P.box.y, P.box.z}, inv[3] = {P.inv_box.x, P.inv_box.y, P.inv_box.z};
    const int K[3] = {P.grid.x, P.grid.y, P.grid.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        p[d] -= floor(p[d] * inv[d] + 0.5) * L[d];
        const double w0 = p[d] * inv[d];
        const double fr = K[d] * (w0 - (int)(w0 + 0.5) + 0.5);
        const int ifr = (int)fr;
        int g = ifr - 5 + 1;
        g += g < 0 ? K[d] : 0;
        g0[d] = g;
        w[d] = fr - ifr;
    }
}

__device__ __forceinline__ void site_splines(const double* pos, int i, const ElecParams& P, int g0[3], double th[3][5][4]) {
    double w[3];
    site_origin(pos, i, P, g0, w);
#pragma unroll
    for (int d = 0; d < 3; ++d) bspline5(w[d], th[d]);
}

__device__ __forceinline__ int wrapg(int v, int K) { return v >= K ? v - K : v; }

// th[k][c] for a lane-dependent k in [0, 5), as a masked sum over the five
// entries: the spline table stays in registers (a dynamic index, or a select
// chain the compiler folds back into one, would move it to scratch memory)
__device__ __forceinline__ double pick5(const double (&th)[5][4], int k, int c) {
    double v = 0.0;
#pragma unroll
    for (int m = 0; m < 5; ++m) v += th[m][c] * (double)(k == m);
    return v;
}

// ---------------------------------------------------------------------------
// PME spreading by grid bricks.  The grid is cut into bricks of B.s points per
// axis, the last brick of an axis taking the remainder (4 to 2 B.s - 1 points,
// never more than 12); a site belongs to the bin of the brick holding its
// origin g0.  Its 5x5x5 support reaches the brick of its own bin and the next
// one along each axis only (every brick is at least 4 points wide), so brick b
// is complete after visiting bins b-1 and b per axis.  One workgroup owns one brick and
// writes every point of it exactly once: no atomics, no clearing pass, and a
// summation order fixed by the (sorted) bin contents.
// ---------------------------------------------------------------------------
// sites whose weight (charge or polarisability) is zero add nothing to the grid and get no bin
__global__ void k_site_bins(int n, const double* __restrict__ pos, const double* __restrict__ weight, ElecParams P,
                            BrickGeom B, int* __restrict__ bin_of, int* __restrict__ bin_count) {
    critical_path_wave();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (weight[i] == 0.0) {
        bin_of[i] = -1;
        return;
    }
    int g0[3];
    double w[3];
    site_origin(pos, i, P, g0, w);
    const int b = (min(g0[0] / B.s[0], B.nb[0] - 1) * B.nb[1] + min(g0[1] / B.s[1], B.nb[1] - 1)) * B.nb[2] +
                  min(g0[2] / B.s[2], B.nb[2] - 1);
    bin_of[i] = b;
    atomicAdd(&bin_count[b], 1);
}

// Per-site spline table for the spreading kernel, computed once per
// evaluation (positions do not change during the dipole iteration):
// g0[3 i + d] = grid origin, sp[30 i + 10 d + k] = value (k < 5) and first
// derivative (k >= 5) of the order-5 B-spline along axis d.
__global__ void k_site_spline_table(int n, const double* __restrict__ pos, ElecParams P, int* __restrict__ g0tab,
                                    double* __restrict__ sptab) {
    critical_path_wave();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int g0[3];
    double th[3][5][4];
    site_splines(pos, i, P, g0, th);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        g0tab[3 * i + d] = g0[d];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            sptab[30 * i + 10 * d + k] = th[d][k][0];
            sptab[30 * i + 10 * d + 5 + k] = th[d][k][1];
        }
    }
}

constexpr int kBrickThreads = 128;
constexpr int kBrickCand = 64;  // candidate sites staged in LDS per pass
constexpr int kBrickZ = 12;     // a brick is at most 12 points wide along every axis

// charges (mode 0) or fractional dipoles (mode 1) onto the real grid (:1556-1620, :2326-2337).
// Each thread owns whole (x, y) columns of the brick (at most two: a brick
// face has at most 144 columns); the candidates' grid origins are broadcast
// from the lanes that staged them, so a column tests a candidate with scalar
// operands and reads LDS only for the candidates that cover it.
__global__ void __launch_bounds__(kBrickThreads) k_spread_bricks(int mode, const double* __restrict__ q,
                                                                 const double* __restrict__ mu, ElecParams P,
                                                                 BrickGeom B, const int* __restrict__ bin_start,
                                                                 const int* __restrict__ members,
                                                                 const int* __restrict__ g0tab,
                                                                 const double* __restrict__ sptab,
                                                                 double* __restrict__ grid) {
    critical_path_wave();
    // per candidate: contribution at support offset (ix, iy, iz) is
    //   u0[iy] v0[iz] A[ix] + T0[ix] (U1[iy] v0[iz] + u0[iy] V1[iz])
    __shared__ double cs[kBrickCand][30];  // A, T0, u0, U1, v0, V1 (5 each)
    __shared__ int cg[kBrickCand][3];
    __shared__ int csite[kBrickCand];
    __shared__ int binlist[8], binoff[9], nbins;
    const int K[3] = {P.grid.x, P.grid.y, P.grid.z};
    const int bz = blockIdx.x % B.nb[2], by = (blockIdx.x / B.nb[2]) % B.nb[1], bx = blockIdx.x / (B.nb[2] * B.nb[1]);
    const int bb[3] = {bx, by, bz};
    int lo[3], sz[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        lo[d] = bb[d] * B.s[d];
        sz[d] = bb[d] == B.nb[d] - 1 ? K[d] - lo[d] : B.s[d];
    }
    if (threadIdx.x == 0) {
        int bl[3][2], nbl[3];
        for (int d = 0; d < 3; ++d) {
            if (B.nb[d] == 1) {
                bl[d][0] = 0;
                nbl[d] = 1;
            } else {
                bl[d][0] = (bb[d] - 1 + B.nb[d]) % B.nb[d];
                bl[d][1] = bb[d];
                nbl[d] = 2;
            }
        }
        int m = 0;
        binoff[0] = 0;
        for (int a = 0; a < nbl[0]; ++a)
            for (int b = 0; b < nbl[1]; ++b)
                for (int c = 0; c < nbl[2]; ++c) {
                    const int id = (bl[0][a] * B.nb[1] + bl[1][b]) * B.nb[2] + bl[2][c];
                    binlist[m] = id;
                    binoff[m + 1] = binoff[m] + (bin_start[id + 1] - bin_start[id]);
                    ++m;
                }
        nbins = m;
    }
    __syncthreads();
    const int total = binoff[nbins];
    const int ncol = sz[0] * sz[1];
    const int t = threadIdx.x;
    const bool has0 = t < ncol, has1 = t + kBrickThreads < ncol;
    const int gx0 = lo[0] + t / sz[1], gy0 = lo[1] + t % sz[1];
    const int t1 = t + kBrickThreads;
    const int gx1 = lo[0] + t1 / sz[1], gy1 = lo[1] + t1 % sz[1];
    const int lane = threadIdx.x & 63;
    const double sc[3] = {K[0] * P.inv_box.x, K[1] * P.inv_box.y, K[2] * P.inv_box.z};
    double acc0[kBrickZ], acc1[kBrickZ];
#pragma unroll
    for (int k = 0; k < kBrickZ; ++k) acc0[k] = acc1[k] = 0.0;

    // add candidate c to the column (gx, gy) if its support covers it
    auto column = [&](int c, int cx, int cy, int cz, int gx, int gy, double (&acc)[kBrickZ]) {
        int ix = gx - cx;
        ix += ix < 0 ? K[0] : 0;
        int iy = gy - cy;
        iy += iy < 0 ? K[1] : 0;
        if (ix < 5 && iy < 5) {
            const double a = cs[c][ix], t0 = cs[c][5 + ix], u0 = cs[c][10 + iy], U1 = cs[c][15 + iy];
            int iz = lo[2] - cz;
            iz += iz < 0 ? K[2] : 0;
#pragma unroll
            for (int k = 0; k < kBrickZ; ++k) {
                // iz walks the column; it leaves [0, 5) for good once it passes 4
                const int z = iz + k >= K[2] ? iz + k - K[2] : iz + k;
                if (k < sz[2] && z < 5) {
                    const double v0 = cs[c][20 + z];
                    acc[k] += u0 * v0 * a + t0 * (U1 * v0 + u0 * cs[c][25 + z]);
                }
            }
        }
    };

    for (int base = 0; base < total; base += kBrickCand) {
        const int nc = min(kBrickCand, total - base);
        if ((int)threadIdx.x < nc) {
            const int c = base + threadIdx.x;
            int m = 0;
            while (c >= binoff[m + 1]) ++m;
            const int site = members[bin_start[binlist[m]] + (c - binoff[m])];
            csite[threadIdx.x] = site;
            cg[threadIdx.x][0] = g0tab[3 * site];
            cg[threadIdx.x][1] = g0tab[3 * site + 1];
            cg[threadIdx.x][2] = g0tab[3 * site + 2];
        }
        __syncthreads();
        // stage the weighted spline rows: element e of candidate c
        for (int e = threadIdx.x; e < nc * 30; e += kBrickThreads) {
            const int c = e / 30, k = e % 30, site = csite[c];
            const double* sp = sptab + 30 * site;
            const int f = k / 5, r = k % 5;
            double v;
            switch (f) {
                case 0: v = mode == 0 ? q[site] * sp[r] : sc[0] * mu[3 * site] * sp[5 + r]; break;  // A
                case 1: v = sp[r]; break;                                                            // T0
                case 2: v = sp[10 + r]; break;                                                       // u0
                case 3: v = mode == 0 ? 0.0 : sc[1] * mu[3 * site + 1] * sp[15 + r]; break;         // U1
                case 4: v = sp[20 + r]; break;                                                       // v0
                default: v = mode == 0 ? 0.0 : sc[2] * mu[3 * site + 2] * sp[25 + r]; break;        // V1
            }
            cs[c][k] = v;
        }
        int ox = 0, oy = 0, oz = 0;
        if (lane < nc) {
            ox = cg[lane][0];
            oy = cg[lane][1];
            oz = cg[lane][2];
        }
        __syncthreads();
        for (int c = 0; c < nc; ++c) {  // uniform loop: the broadcasts run with every lane active
            const int cx = __builtin_amdgcn_readlane(ox, c), cy = __builtin_amdgcn_readlane(oy, c),
                      cz = __builtin_amdgcn_readlane(oz, c);
            if (has0) column(c, cx, cy, cz, gx0, gy0, acc0);
            if (has1) column(c, cx, cy, cz, gx1, gy1, acc1);
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < kBrickZ; ++k) {
        if (k < sz[2]) {
            if (has0) grid[((size_t)gx0 * K[1] + gy0) * K[2] + lo[2] + k] = acc0[k];
            if (has1) grid[((size_t)gx1 * K[1] + gy1) * K[2] + lo[2] + k] = acc1[k];
        }
    }
}

// ---------------------------------------------------------------------------
// Dipole spreading as a sparse product.  With the positions fixed for a whole
// evaluation, the fractional-dipole grid of every iteration is linear in the
// dipoles: grid[g] = sum over the (site, support point) entries of g of
// mu_site . c, c = (sc_x t1 u0 v0, sc_y t0 u1 v0, sc_z t0 u0 v1).  The entries
// are tabulated once per evaluation in grid-point order (rows sorted by site,
// so every sum has a fixed order); each iteration then streams the table.
// ---------------------------------------------------------------------------
__device__ __forceinline__ size_t support_point(const int* g0, int k, const ElecParams& P, int& a, int& b, int& c) {
    a = k / 25;
    b = (k / 5) % 5;
    c = k % 5;
    const int gx = wrapg(g0[0] + a, P.grid.x), gy = wrapg(g0[1] + b, P.grid.y), gz = wrapg(g0[2] + c, P.grid.z);
    return ((size_t)gx * P.grid.y + gy) * P.grid.z + gz;
}

__global__ void k_csr_count(int npol, const int* __restrict__ polidx, const int* __restrict__ g0tab, ElecParams P,
                            int* __restrict__ rowcount) {
    critical_path_wave();
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)npol * 125) return;
    const int s = polidx[t / 125];
    const int g0[3] = {g0tab[3 * s], g0tab[3 * s + 1], g0tab[3 * s + 2]};
    int a, b, c;
    atomicAdd(&rowcount[support_point(g0, (int)(t % 125), P, a, b, c)], 1);
}

__global__ void k_csr_fill(int npol, const int* __restrict__ polidx, const int* __restrict__ g0tab,
                           const double* __restrict__ sptab, ElecParams P, const int* __restrict__ rowstart,
                           int* __restrict__ fill, int* __restrict__ esite, double* __restrict__ ecoef) {
    critical_path_wave();
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)npol * 125) return;
    const int s = polidx[t / 125];
    const int g0[3] = {g0tab[3 * s], g0tab[3 * s + 1], g0tab[3 * s + 2]};
    int a, b, c;
    const size_t g = support_point(g0, (int)(t % 125), P, a, b, c);
    const int e = rowstart[g] + atomicAdd(&fill[g], 1);
    const double* sp = sptab + 30 * (size_t)s;
    const double t0 = sp[a], t1 = sp[5 + a], u0 = sp[10 + b], u1 = sp[15 + b], v0 = sp[20 + c], v1 = sp[25 + c];
    esite[e] = s;
    ecoef[3 * (size_t)e] = P.grid.x * P.inv_box.x * (t1 * u0 * v0);
    ecoef[3 * (size_t)e + 1] = P.grid.y * P.inv_box.y * (t0 * u1 * v0);
    ecoef[3 * (size_t)e + 2] = P.grid.z * P.inv_box.z * (t0 * u0 * v1);
}

// rows ascending by site (a row holds a handful of entries)
__global__ void k_csr_sort(size_t ngrid, const int* __restrict__ rowstart, int* __restrict__ esite,
                           double* __restrict__ ecoef) {
    critical_path_wave();
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngrid) return;
    const int b = rowstart[g], e = rowstart[g + 1];
    for (int k = b + 1; k < e; ++k) {
        const int v = esite[k];
        const double c0 = ecoef[3 * (size_t)k], c1 = ecoef[3 * (size_t)k + 1], c2 = ecoef[3 * (size_t)k + 2];
        int m = k - 1;
        while (m >= b && esite[m] > v) {
            esite[m + 1] = esite[m];
            ecoef[3 * (size_t)(m + 1)] = ecoef[3 * (size_t)m];
            ecoef[3 * (size_t)(m + 1) + 1] = ecoef[3 * (size_t)m + 1];
            ecoef[3 * (size_t)(m + 1) + 2] = ecoef[3 * (size_t)m + 2];
            --m;
        }
        esite[m + 1] = v;
        ecoef[3 * (size_t)(m + 1)] = c0;
        ecoef[3 * (size_t)(m + 1) + 1] = c1;
        ecoef[3 * (size_t)(m + 1) + 2] = c2;
    }
}

__global__ void __launch_bounds__(256) k_spread_csr(size_t ngrid, const int* __restrict__ rowstart,
                                                    const int* __restrict__ esite, const double* __restrict__ ecoef,
                                                    const double* __restrict__ mu, double* __restrict__ grid,
                                                    const int* __restrict__ stop) {
    critical_path_wave();
    if (stop && *stop) return;
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngrid) return;
    double v = 0.0;
    for (int e = rowstart[g]; e < rowstart[g + 1]; ++e) {
        const int s = esite[e];
        v += mu[3 * s] * ecoef[3 * (size_t)e] + mu[3 * s + 1] * ecoef[3 * (size_t)e + 1] +
             mu[3 * s + 2] * ecoef[3 * (size_t)e + 2];
    }
    grid[g] = v;
}


// multiply the half-complex spectrum by the reciprocal-space influence function (:1676-1713)
__global__ void k_convolve(size_t nhalf, const double* __restrict__ influence, double2* __restrict__ spec) {
    critical_path_wave();
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nhalf) return;
    const double e = influence[t];
    double2 v = spec[t];
    v.x *= e;
    v.y *= e;
    spec[t] = v;
}

// influence function on the half grid (kz <= K3/2)
__global__ void k_influence(ElecParams P, int khp, const double* __restrict__ bx, const double* __restrict__ by,
                            const double* __restrict__ bz, double* __restrict__ out) {
    critical_path_wave();
    const size_t total = (size_t)P.grid.x * P.grid.y * khp;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int kz = (int)(t % khp), ky = (int)((t / khp) % P.grid.y), kx = (int)(t / ((size_t)khp * P.grid.y));
    if ((kx == 0 && ky == 0 && kz == 0) || kz >= P.grid.z / 2 + 1) {
        out[t] = 0.0;
        return;
    }
    const double expf_ = (M_PI * M_PI) / (P.alpha * P.alpha);
    const double sf = 1.0 / (M_PI * P.box.x * P.box.y * P.box.z);
    const int mx = (kx < (P.grid.x + 1) / 2) ? kx : (kx - P.grid.x);
    const int my = (ky < (P.grid.y + 1) / 2) ? ky : (ky - P.grid.y);
    const int mz = (kz < (P.grid.z + 1) / 2) ? kz : (kz - P.grid.z);
    const double hx = mx * P.inv_box.x, hy = my * P.inv_box.y, hz = mz * P.inv_box.z;
    const double m2 = hx * hx + hy * hy + hz * hz;
    const double den = m2 * bx[kx] * by[ky] * bz[kz];
    out[t] = sf * exp(-expf_ * m2) / den;
}

// potential and derivatives to second order at each site (10 components,
// order 000 100 010 001 200 020 002 110 101 011), :1715-1821
__global__ void __launch_bounds__(64) k_gather(int n, const double* __restrict__ pos, ElecParams P, const double* __restrict__ grid,
                         double* __restrict__ phi) {
    critical_path_wave();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int g0[3];
    double th[3][5][4];
    site_splines(pos, i, P, g0, th);
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int iz = 0; iz < 5; ++iz) {
        const int k = wrapg(g0[2] + iz, P.grid.z);
        const double v0 = th[2][iz][0], v1 = th[2][iz][1], v2 = th[2][iz][2];
        double tu00 = 0, tu10 = 0, tu01 = 0, tu20 = 0, tu11 = 0, tu02 = 0;
#pragma unroll
        for (int iy = 0; iy < 5; ++iy) {
            const int j = wrapg(g0[1] + iy, P.grid.y);
            const double u0 = th[1][iy][0], u1 = th[1][iy][1], u2 = th[1][iy][2];
            double t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
            for (int ix = 0; ix < 5; ++ix) {
                const int ii = wrapg(g0[0] + ix, P.grid.x);
                const double gv = grid[((size_t)ii * P.grid.y + j) * P.grid.z + k];
                t0 += gv * th[0][ix][0];
                t1 += gv * th[0][ix][1];
                t2 += gv * th[0][ix][2];
            }
            tu00 += t0 * u0;
            tu10 += t1 * u0;
            tu01 += t0 * u1;
            tu20 += t2 * u0;
            tu11 += t1 * u1;
            tu02 += t0 * u2;
        }
        acc[0] += tu00 * v0;
        acc[1] += tu10 * v0;
        acc[2] += tu01 * v0;
        acc[3] += tu00 * v1;
        acc[4] += tu20 * v0;
        acc[5] += tu02 * v0;
        acc[6] += tu00 * v2;
        acc[7] += tu11 * v0;
        acc[8] += tu10 * v1;
        acc[9] += tu01 * v1;
    }
#pragma unroll
    for (int c = 0; c < 10; ++c) phi[10 * i + c] = acc[c];
}

// field = -scale * grad(phi) (+= for the induced field, = for the fixed field); :2267-2278, :2300-2324
__global__ void k_record_field(int n, ElecParams P, const double* __restrict__ phi, int accumulate, double self_term,
                               const double* __restrict__ mu, double* __restrict__ field) {
    critical_path_wave();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double sc[3] = {P.grid.x * P.inv_box.x, P.grid.y * P.inv_box.y, P.grid.z * P.inv_box.z};
    for (int d = 0; d < 3; ++d) {
        double f = -sc[d] * phi[10 * i + 1 + d];
        if (mu) f += mu[3 * i + d] * self_term;
        field[3 * i + d] = accumulate ? field[3 * i + d] + f : f;
    }
}


// ---------------------------------------------------------------------------
// One induced-dipole iteration for every polarisable site (convergeInduceDipoles
// :557-616, calculateInducedDipoleFields :438-486 / :2300-2390, updateInducedDipole
// :516-531): real-space field from the P row and the prefactors, PME gradient of
// the dipole potential (4 of the 10 gather components, one lane per (ix, iy)
// column), self term, SOR step mu_out = mu_in + omega (alpha (E_fix + E_ind) - mu_in)
// and the per-block convergence sums.  mu_in and mu_out are distinct buffers,
// so every field is taken from the dipoles before the step, as in the reference.
// ---------------------------------------------------------------------------
// Field at polarisable site i from the dipoles v, summed over the G lanes of
// its group: real space over the P row with the prefactors, and with PME minus
// the gradient of the reciprocal potential on the grid (4 of the 10 gather
// components, one lane per (ix, iy) column, lanes 0..24).  The self term is
// left to the caller.
template <int G>
__device__ __forceinline__ void site_field(int i, int lane, const double* __restrict__ pos, const ElecParams& P,
                                           int capP, const int* __restrict__ nbP, const int* __restrict__ cntP,
                                           const double2* __restrict__ pf, const double* __restrict__ v,
                                           const double* __restrict__ grid, double& fx, double& fy, double& fz) {
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    double ex = 0, ey = 0, ez = 0;
    const int c = min(cntP[i], capP);
    const int* row = nbP + (size_t)i * capP;
    const double2* prow = pf + (size_t)i * capP;
    for (int k = lane; k < c; k += G) {
        const int j = row[k];
        double dx = pos[3 * j] - xi, dy = pos[3 * j + 1] - yi, dz = pos[3 * j + 2] - zi;
        min_image(dx, dy, dz, P);
        const double2 p = prow[k];
        const double mx = v[3 * j], my = v[3 * j + 1], mz = v[3 * j + 2];
        const double dur = (mx * dx + my * dy + mz * dz) * p.y;
        ex += dx * dur + mx * p.x;
        ey += dy * dur + my * p.x;
        ez += dz * dur + mz * p.x;
    }
    if (P.pme && lane < 25) {
        int g0[3];
        double th[3][5][4];
        site_splines(pos, i, P, g0, th);
        const int a = lane / 5, b = lane % 5;
        const int ii = wrapg(g0[0] + a, P.grid.x), jj = wrapg(g0[1] + b, P.grid.y);
        const double* line = grid + ((size_t)ii * P.grid.y + jj) * P.grid.z;
        double t0 = 0, t1 = 0;
#pragma unroll
        for (int iz = 0; iz < 5; ++iz) {
            const double gv = line[wrapg(g0[2] + iz, P.grid.z)];
            t0 += gv * th[2][iz][0];
            t1 += gv * th[2][iz][1];
        }
        const double x0 = pick5(th[0], a, 0), x1 = pick5(th[0], a, 1), y0 = pick5(th[1], b, 0), y1 = pick5(th[1], b, 1);
        ex -= P.grid.x * P.inv_box.x * (x1 * y0 * t0);
        ey -= P.grid.y * P.inv_box.y * (x0 * y1 * t0);
        ez -= P.grid.z * P.inv_box.z * (x0 * y0 * t1);
    }
    fx = group_sum<G>(ex);
    fy = group_sum<G>(ey);
    fz = group_sum<G>(ez);
}

// per-block sum of one value per group (lane 0 of each group holds it)
template <int G>
__device__ __forceinline__ void block_partial(double v, int lane, double* __restrict__ out) {
    __shared__ double red[256 / G];
    if (lane == 0) red[threadIdx.x / G] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < 256 / G; ++k) s += red[k];
        *out = s;
    }
    __syncthreads();  // red may be reused by the next call
}

// ---------------------------------------------------------------------------
// one SOR step for the polarisable sites (convergeInduceDipoles and
// calculateInducedDipoleField, MBPolReferenceElectrostaticsForce.cpp:557-616,
// :516-531): field of mu_in, self term, SOR step
// mu_out = mu_in + omega (alpha (E_fix + E_ind) - mu_in) and the per-block
// convergence sums.  mu_in and mu_out are distinct buffers, so every field is
// taken from the dipoles before the step, as in the reference.
// ---------------------------------------------------------------------------
template <int G>
__global__ void __launch_bounds__(256) k_iterate(int npol, const int* __restrict__ polidx, const double* __restrict__ pos,
                                                 ElecParams P, int capP, const int* __restrict__ nbP,
                                                 const int* __restrict__ cntP, const double2* __restrict__ pf,
                                                 const double* __restrict__ mu_in, double* __restrict__ mu_out,
                                                 const double* __restrict__ efix_pol, const double* __restrict__ pol,
                                                 const double* __restrict__ grid, double self_term, double omega,
                                                 double* __restrict__ partial, const SolveCtl* __restrict__ ctl) {
    critical_path_wave();
    if (ctl->stop) return;
    const int g = (blockIdx.x * blockDim.x + threadIdx.x) / G, lane = threadIdx.x % G;
    double e2 = 0.0;
    if (g < npol) {
        const int i = polidx[g];
        double ex, ey, ez;
        site_field<G>(i, lane, pos, P, capP, nbP, cntP, pf, mu_in, grid, ex, ey, ez);
        if (lane == 0) {
            const double m[3] = {mu_in[3 * i], mu_in[3 * i + 1], mu_in[3 * i + 2]};
            const double f[3] = {ex + m[0] * self_term, ey + m[1] * self_term, ez + m[2] * self_term};
            const double a = pol[i];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double delta = efix_pol[3 * i + d] + f[d] * a - m[d];
                mu_out[3 * i + d] = m[d] + delta * omega;
                e2 += delta * delta;
            }
        }
    }
    block_partial<G>(e2, lane, partial + blockIdx.x);
}

// ---------------------------------------------------------------------------
// Preconditioned conjugate gradients for the same linear system, written in
// its symmetric form A mu = E_fix with A = diag(1/alpha) - T, where T mu is
// the induced field (real space, reciprocal and self term) that the SOR step
// uses; the preconditioner is the diagonal of A, 1/alpha - self.  At every
// step the solver also forms the reference's convergence measure of the
// current dipoles, omega * debye * rms(alpha (E_fix + T mu) - mu) = omega *
// debye * rms(alpha r), so it stops by the same rule as the SOR iteration.
// ---------------------------------------------------------------------------
// mode 0: r = E_fix - A v (the initial residual); mode 1: q = A v, partials (v.q)
template <int G>
__global__ void __launch_bounds__(256) k_cg_matvec(int npol, const int* __restrict__ polidx, const double* __restrict__ pos,
                                                   ElecParams P, int capP, const int* __restrict__ nbP,
                                                   const int* __restrict__ cntP, const double2* __restrict__ pf,
                                                   const double* __restrict__ v, const double* __restrict__ grid,
                                                   double self_term, const double* __restrict__ pol,
                                                   const double* __restrict__ efix, int mode, double* __restrict__ out,
                                                   double* __restrict__ partial, const SolveCtl* __restrict__ ctl) {
    critical_path_wave();
    if (ctl->stop) return;
    const int g = (blockIdx.x * blockDim.x + threadIdx.x) / G, lane = threadIdx.x % G;
    double s0 = 0.0, s1 = 0.0;
    if (g < npol) {
        const int i = polidx[g];
        double ex, ey, ez;
        site_field<G>(i, lane, pos, P, capP, nbP, cntP, pf, v, grid, ex, ey, ez);
        if (lane == 0) {
            const double ia = 1.0 / pol[i];
            const double f[3] = {ex, ey, ez};
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double vd = v[3 * i + d];
                const double av = vd * ia - (f[d] + vd * self_term);  // (A v)_d
                if (mode == 0) {
                    out[3 * i + d] = efix[3 * i + d] - av;
                } else {
                    out[3 * i + d] = av;
                    s0 += vd * av;
                }
            }
        }
    }
    block_partial<G>(s0, lane, partial + 2 * blockIdx.x);
    block_partial<G>(s1, lane, partial + 2 * blockIdx.x + 1);
}

// Block-Jacobi preconditioner: the polarisable sites of one molecule (at most
// kBlockSites, larger molecules split into several blocks) form a block of
// A restricted to the molecule, taken as the isolated molecule's polarisation
// matrix diag(1/alpha) - T_Thole (the Thole-damped dipole tensor of the
// molecule's own pairs, without the Ewald split: in PME the real-space and
// reciprocal parts of an intramolecular pair add up to it, and the self term
// cancels the reciprocal self-interaction).  One thread inverts one block in
// place (Gauss-Jordan; the block is symmetric positive definite).
constexpr int kBlockSites = 4, kBlockLd = 3 * kBlockSites;
constexpr int kGB = 16;  // lanes per block when applying the preconditioner (>= kBlockLd)

// kGB lanes per block, lane l < 3m holding row l of the block (site l / 3,
// component l % 3) in registers: assembly, then in-place Gauss-Jordan with the
// pivot row broadcast by shuffles.  The stored inverse is kBlockLd x kBlockLd,
// zero outside the block's 3m x 3m.
__global__ void __launch_bounds__(256) k_block_factor(int nblocks, const int* __restrict__ blk_start,
                                                      const int* __restrict__ blk_site, const double* __restrict__ pos,
                                                      const SiteParams S, ElecParams P, double* __restrict__ minv) {
    critical_path_wave();
    const int b = (blockIdx.x * blockDim.x + threadIdx.x) / kGB, l = threadIdx.x % kGB;
    if (b >= nblocks) return;  // uniform over the lanes of a block
    const int s0 = blk_start[b], m = blk_start[b + 1] - s0, d = 3 * m;
    double a[kBlockLd];
#pragma unroll
    for (int c = 0; c < kBlockLd; ++c) a[c] = 0.0;
    if (l < d) {
        const int u = l / 3, x = l % 3, i = blk_site[s0 + u];
#pragma unroll
        for (int w = 0; w < kBlockSites; ++w) {
            if (w >= m) continue;
            if (w == u) {
                const double ia = 1.0 / S.pol[i];
#pragma unroll
                for (int y = 0; y < 3; ++y)
                    if (y == x) a[3 * w + y] = ia;
                continue;
            }
            const int j = blk_site[s0 + w];
            double dx = pos[3 * j] - pos[3 * i], dy = pos[3 * j + 1] - pos[3 * i + 1],
                   dz = pos[3 * j + 2] - pos[3 * i + 2];
            min_image(dx, dy, dz, P);
            const double r2 = dx * dx + dy * dy + dz * dz, r = sqrt(r2);
            const Thole th(S.damp6[i], S.damp6[j], r);
            const double g = tdd_gamma(P, 1, S.type[i], S.type[j]);
            const double e = th.ex(g);
            const double p1 = th.s3(e) / (r2 * r), p2 = -3.0 * th.s5(g, e) / (r2 * r2 * r);
            const double dv[3] = {dx, dy, dz};
            const double dxl = x == 0 ? dx : (x == 1 ? dy : dz);
#pragma unroll
            for (int y = 0; y < 3; ++y) a[3 * w + y] = (y == x ? p1 : 0.0) + p2 * dxl * dv[y];
        }
    }
    for (int k = 0; k < d; ++k) {  // d is uniform over the group
        double rk[kBlockLd];
#pragma unroll
        for (int c = 0; c < kBlockLd; ++c) rk[c] = __shfl(a[c], k, kGB);
        double akk = 0.0, aik = 0.0;
#pragma unroll
        for (int c = 0; c < kBlockLd; ++c)
            if (c == k) {
                akk = rk[c];
                aik = a[c];
            }
        const double piv = 1.0 / akk;
#pragma unroll
        for (int c = 0; c < kBlockLd; ++c) rk[c] = (c == k ? 1.0 : rk[c]) * piv;
        if (l == k) {
#pragma unroll
            for (int c = 0; c < kBlockLd; ++c) a[c] = rk[c];
        } else {
#pragma unroll
            for (int c = 0; c < kBlockLd; ++c) a[c] = (c == k ? 0.0 : a[c]) - aik * rk[c];
        }
    }
    if (l < kBlockLd) {
        double* M = minv + ((size_t)b * kBlockLd + l) * kBlockLd;
#pragma unroll
        for (int c = 0; c < kBlockLd; ++c) M[c] = (l < d && c < d) ? a[c] : 0.0;
    }
}

// z = M^-1 r on one block with kGB lanes, lane l < 3m owning row l (site l / 3,
// component l % 3, global index t); rl is this lane's r (0 beyond the block)
__device__ __forceinline__ double block_row_apply(int b, int lane, double rl, const double* __restrict__ minv) {
    const double* M = minv + (size_t)b * kBlockLd * kBlockLd + (size_t)min(lane, kBlockLd - 1) * kBlockLd;
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < kBlockLd; ++c) acc += M[c] * __shfl(rl, c, kGB);
    return acc;
}

// the initial residual's preconditioned vector and partials (r.z, |alpha r|^2)
__global__ void __launch_bounds__(256) k_cg_precond(int nblocks, const int* __restrict__ blk_start,
                                                    const int* __restrict__ blk_site, const double* __restrict__ minv,
                                                    const double* __restrict__ pol, const double* __restrict__ r,
                                                    double* __restrict__ z, double* __restrict__ partial,
                                                    double* __restrict__ blk_ee, const SolveCtl* __restrict__ ctl) {
    critical_path_wave();
    if (ctl->stop) return;
    const int b = (blockIdx.x * blockDim.x + threadIdx.x) / kGB, lane = threadIdx.x % kGB;
    double s0 = 0.0, s1 = 0.0;
    if (b < nblocks) {  // uniform over the lanes of a block
        const int st = blk_start[b], d = 3 * (blk_start[b + 1] - st);
        int t = 0;
        double rl = 0.0, a = 0.0;
        if (lane < d) {
            const int i = blk_site[st + lane / 3];
            t = 3 * i + lane % 3;
            rl = r[t];
            a = pol[i];
        }
        const double zl = block_row_apply(b, lane, rl, minv);
        if (lane < d) {
            z[t] = zl;
            s0 = rl * zl;
            s1 = (a * rl) * (a * rl);
        }
        if (blk_ee) {  // per block, for the per-group convergence measure
            double e = s1;
#pragma unroll
            for (int off = kGB / 2; off > 0; off >>= 1) e += __shfl_xor(e, off, kGB);
            if (lane == 0) blk_ee[b] = e;
        }
    }
    block_partial<1>(s0, 0, partial + 2 * blockIdx.x);
    block_partial<1>(s1, 0, partial + 2 * blockIdx.x + 1);
}

// sum of the entries 2b + c (b < nb) of partial, in an order fixed by nb
// alone, so that every block of a launch obtains the same value
__device__ __forceinline__ double reduce_partials(int nb, int c, const double* __restrict__ partial) {
    __shared__ double red[256];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) s += partial[2 * b + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// step k: a = rz[k] / (p.q); x += a p; r -= a q; z = M^-1 r; partials (r.z, |alpha r|^2);
// kGB lanes per preconditioner block
__global__ void __launch_bounds__(256) k_cg_update(int nblocks, const int* __restrict__ blk_start,
                                                   const int* __restrict__ blk_site, const double* __restrict__ minv,
                                                   const double* __restrict__ pol, int nb_mv,
                                                   const double* __restrict__ partial_mv,
                                                   const double* __restrict__ rz, int k, double* __restrict__ x,
                                                   double* __restrict__ r, const double* __restrict__ p,
                                                   const double* __restrict__ q, double* __restrict__ z,
                                                   double* __restrict__ partial, double* __restrict__ blk_ee,
                                                   const SolveCtl* __restrict__ ctl) {
    critical_path_wave();
    if (ctl->stop) return;
    const double pq = reduce_partials(nb_mv, 0, partial_mv);
    const double step = rz[k] / pq;
    const int b = (blockIdx.x * blockDim.x + threadIdx.x) / kGB, lane = threadIdx.x % kGB;
    double s0 = 0.0, s1 = 0.0;
    if (b < nblocks) {  // uniform over the lanes of a block
        const int st = blk_start[b], d = 3 * (blk_start[b + 1] - st);
        int t = 0;
        double rl = 0.0, a = 0.0;
        if (lane < d) {
            const int i = blk_site[st + lane / 3];
            t = 3 * i + lane % 3;
            x[t] += step * p[t];
            rl = r[t] - step * q[t];
            r[t] = rl;
            a = pol[i];
        }
        const double zl = block_row_apply(b, lane, rl, minv);
        if (lane < d) {
            z[t] = zl;
            s0 = rl * zl;
            s1 = (a * rl) * (a * rl);
        }
        if (blk_ee) {  // per block, for the per-group convergence measure
            double e = s1;
#pragma unroll
            for (int off = kGB / 2; off > 0; off >>= 1) e += __shfl_xor(e, off, kGB);
            if (lane == 0) blk_ee[b] = e;
        }
    }
    block_partial<1>(s0, 0, partial + 2 * blockIdx.x);
    block_partial<1>(s1, 0, partial + 2 * blockIdx.x + 1);
}

// k-th direction: rz[k] = r.z; p = z (k == 0) or z + (rz[k] / rz[k-1]) p;
// the measure omega * debye * rms(alpha r) over all n sites of the current
// dipoles (k + 1 field evaluations so far) and the stopping rules: converged
// below tol, not converged after max_iter + 1 field evaluations
__global__ void __launch_bounds__(256) k_cg_direction(int npol, const int* __restrict__ polidx, int nb,
                                                      const double* __restrict__ partial, double* __restrict__ rz,
                                                      int k, const double* __restrict__ z, double* __restrict__ p,
                                                      int n_all, double tol, int max_iter, GroupConv gc,
                                                      SolveCtl* __restrict__ ctl) {
    critical_path_wave();
    if (ctl->stop) return;
    const double s_rz = reduce_partials(nb, 0, partial);
    const double s_ee = reduce_partials(nb, 1, partial);
    // independent groups: the measure of the worst group (block 0 only needs it)
    double g_eps = 0.0;
    if (gc.ngroups > 0 && blockIdx.x == 0) {
        __shared__ double gmax[256];
        double m = 0.0;
        for (int g = threadIdx.x; g < gc.ngroups; g += 256) {
            double e = 0.0;
            for (int b = gc.blk_start[g]; b < gc.blk_start[g + 1]; ++b) e += gc.blk_ee[b];
            m = fmax(m, el_polar_sor * el_debye * sqrt(e / (double)gc.nsites[g]));
        }
        gmax[threadIdx.x] = m;
        __syncthreads();
        for (int k = 128; k > 0; k >>= 1) {
            if (threadIdx.x < k) gmax[threadIdx.x] = fmax(gmax[threadIdx.x], gmax[threadIdx.x + k]);
            __syncthreads();
        }
        g_eps = gmax[0];
    }
    const double beta = k == 0 ? 0.0 : s_rz / rz[k - 1];
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < npol) {
        const int i = polidx[g];
        // the first direction is z itself: p is not read (a fresh buffer may hold NaN, and 0 * NaN is NaN)
#pragma unroll
        for (int d = 0; d < 3; ++d) p[3 * i + d] = k == 0 ? z[3 * i + d] : z[3 * i + d] + beta * p[3 * i + d];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        rz[k] = s_rz;
        const double eps = gc.ngroups > 0 ? g_eps : el_polar_sor * el_debye * sqrt(s_ee / (double)n_all);
        ctl->eps = eps;
        ctl->iterations = k + 1;
        if (eps < tol) {
            ctl->converged = 1;
            ctl->stop = 1;
        } else if (k + 1 > max_iter) {
            ctl->stop = 1;
        }
    }
}

// SOR step t: epsilon = omega * debye * sqrt(sum / n) from the block partials,
// in block order (:504-511, :588), and the reference's stopping rules
// (:596-612): converged below tol; stop without convergence when epsilon grew
// or after max_iter + 1 steps
__global__ void k_sor_finish(int nblocks, int n, const double* __restrict__ partial, int t, double tol, int max_iter,
                             SolveCtl* __restrict__ ctl) {
    critical_path_wave();
    if (ctl->stop) return;
    __shared__ double red[256];
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += partial[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double eps = el_polar_sor * el_debye * sqrt(red[0] / (double)n);
        const double cur = t == 0 ? 1.0e+50 : ctl->eps;
        ctl->eps = eps;
        ctl->iterations = t + 1;
        if (eps < tol) {
            ctl->converged = 1;
            ctl->stop = 1;
        } else if (cur < eps || t >= max_iter) {
            ctl->stop = 1;
        }
    }
}

__global__ void k_scale_copy(int n3, const double* __restrict__ a, const double* __restrict__ pol, double* __restrict__ out,
                             double* __restrict__ out2) {
    critical_path_wave();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n3) return;
    const double v = a[t] * pol[t / 3];
    out[t] = v;
    if (out2) out2[t] = v;
}

// ---------------------------------------------------------------------------
// energy, forces and site potentials (direct space)
// PME: calculatePmeDirectElectrostaticPairIxn (:2510-2716)
// no cutoff: calculateElectrostaticPairIxn (:649-836), with the charge-derivative
// terms (:791-827) gathered as a potential V (same sums, see the driver)
// Each directed pair (i, j) adds its share to site i only; energy is halved.
// ---------------------------------------------------------------------------
template <int G, bool PME>  // PME: Ewald real-space terms (compile-time, so only one branch holds registers)
__global__ void __launch_bounds__(256, 3) k_elec_energy(int n, const double* __restrict__ pos, const SiteParams S,
                                                     ElecParams P, int capP, const int* __restrict__ nbP,
                                                     const int* __restrict__ cntP, int capQ,
                                                     const int* __restrict__ nbQ, const int* __restrict__ cntQ,
                                                     const double* __restrict__ mu, double* __restrict__ force,
                                                     double* __restrict__ V, double* __restrict__ eblk,
                                                     double* __restrict__ esite) {
    critical_path_wave();
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / G, lane = threadIdx.x % G;
    double en = 0, fx = 0, fy = 0, fz = 0, vi = 0;
    if (i < n) {
        // the potential at i is used only at the charged sites of a water (4w+1..4w+3)
        // when the charges depend on the geometry (k_charge_deriv_forces)
        const bool vneed = P.redistribute && (i & 3) != 0;
        const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
        const double ci = S.q[i];
        const double mix = mu[3 * i], miy = mu[3 * i + 1], miz = mu[3 * i + 2];
        for_rows<G>(i, lane, capP, nbP, cntP, capQ, nbQ, cntQ, [&](int j) {
            double dx = pos[3 * j] - xi, dy = pos[3 * j + 1] - yi, dz = pos[3 * j + 2] - zi;
            min_image(dx, dy, dz, P);
            const double r2 = dx * dx + dy * dy + dz * dz, r = sqrt(r2);
            const double ck = S.q[j];
            const double mjx = mu[3 * j], mjy = mu[3 * j + 1], mjz = mu[3 * j + 2];
            const int same = S.mol[i] == S.mol[j];
            const Thole th(S.damp6[i], S.damp6[j], r);
            const double gdd = tdd_gamma(P, same, S.type[i], S.type[j]);
            const double rr1 = 1.0 / r, r2i = rr1 * rr1, rr3 = rr1 * r2i, rr5 = 3.0 * rr3 * r2i,
                         rr7 = 5.0 * rr5 * r2i;
            const double sci3 = mix * dx + miy * dy + miz * dz;  // mu_i . d
            const double sci4 = mjx * dx + mjy * dy + mjz * dz;  // mu_j . d
            const double mm = mix * mjx + miy * mjy + miz * mjz;
            const double gl0 = ci * ck;
            const double gli1 = ck * sci3 - ci * sci4;
            // the charge-charge (TCC) and charge-dipole (TCD) scales of a pair
            // within one molecule are never used.  The charge-charge scales enter
            // only multiplied by the charge product q_i q_j, or the potential at i
            // (s1: times q_j; direct sum, s3: times the dipole of j); where those
            // factors are zero (an uncharged site, a site whose potential is not
            // used) their exponential, and the incomplete gamma function of s1,
            // are skipped and the zero products stay zero.
            const bool need_s1 = !same && ck != 0.0 && (ci != 0.0 || vneed);
            const bool need_cc = PME ? need_s1 : (need_s1 || (!same && vneed));
            const double ecc = need_cc ? th.ex(P.thole[0]) : 0., ecd = same ? 0. : th.ex(P.thole[1]);
            const double edd = th.ex(gdd);
            const double s5DD = th.s5(gdd, edd), s7DD = th.s7(gdd, edd);
            double gx, gy, gz;  // (ftm2 + ftm2i): force on i is -f * this
            if constexpr (PME) {
                const Ewald ew(P.alpha, r);
                double s1CC = 0., s3CD = 0., s3CC = 0., s5CD = 0.;
                if (!same) {
                    s1CC = need_s1 ? th.s1(P.thole[0], P.tcc_q, ecc) : 0.;
                    s3CD = th.s3(ecd);
                    s3CC = need_cc ? th.s3(ecc) : 0.;
                    s5CD = th.s5(P.thole[1], ecd);
                }
                const double e = ew.bn0 * gl0 - rr1 * gl0 * (1 - s1CC);
                const double ei = 0.5 * (ew.bn1 * gli1) - 0.5 * (rr3 * gli1 * (1 - s3CD));
                en += e + ei;
                vi += ck * (ew.bn0 - rr1 * (1 - s1CC)) - sci4 * (ew.bn1 - rr3 * (1 - s3CD));
                const double gf1 = ew.bn1 * gl0 - (1 - s3CC) * rr3 * gl0;
                const double gfi1 = ew.bn2 * (gli1 + mm) - ew.bn3 * sci3 * sci4;
                const double gfri1 = rr5 * (gli1 * (1 - s5CD) + mm * (1 - s5DD)) - rr7 * sci3 * sci4 * (1 - s7DD);
                const double a = gf1 + gfi1 - gfri1;
                const double bi = -ck * ew.bn1 + sci4 * ew.bn2 - sci4 * rr5 * (1 - s5DD) + ck * rr3 * (1 - s3CD);
                const double bj = ci * ew.bn1 + sci3 * ew.bn2 - sci3 * rr5 * (1 - s5DD) - ci * rr3 * (1 - s3CD);
                gx = dx * a + mix * bi + mjx * bj;
                gy = dy * a + miy * bi + mjy * bj;
                gz = dz * a + miz * bi + mjz * bj;
            } else {
                double bi = sci4 * rr5 * s5DD, bj = sci3 * rr5 * s5DD;
                double a = rr5 * mm * s5DD - rr7 * sci3 * sci4 * s7DD;
                if (!same) {
                    const double s1CC = need_s1 ? th.s1(P.thole[0], P.tcc_q, ecc) : 0., s3CD = th.s3(ecd);
                    const double s3CC = need_cc ? th.s3(ecc) : 0., s5CD = th.s5(P.thole[1], ecd);
                    en += rr1 * gl0 * s1CC + 0.5 * (rr3 * gli1 * s3CD);
                    a += rr3 * gl0 * s3CC + rr5 * gli1 * s5CD;
                    bi += -ck * rr3 * s3CD;
                    bj += ci * rr3 * s3CD;
                    if (P.redistribute) {  // potential at i from the charge and dipole of j (TCC damping)
                        vi += ck * rr1 * s1CC + s3CC * (-sci4) * rr3;
                    }
                }
                gx = dx * a + mix * bi + mjx * bj;
                gy = dy * a + miy * bi + mjy * bj;
                gz = dz * a + miz * bi + mjz * bj;
            }
            fx -= gx;
            fy -= gy;
            fz -= gz;
        });
        fx = group_sum<G>(fx);
        fy = group_sum<G>(fy);
        fz = group_sum<G>(fz);
        vi = group_sum<G>(vi);
        if (esite) {
            const double es = group_sum<G>(en);
            if (lane == 0) esite[i] = 0.5 * es * el_electric;
        }
        if (lane == 0) {
            const double f = el_electric;
            force[3 * i] += fx * f;
            force[3 * i + 1] += fy * f;
            force[3 * i + 2] += fz * f;
            V[i] += vi;
        }
    }
    // energy: each pair seen twice
    __shared__ double red[256];
    red[threadIdx.x] = 0.5 * en * el_electric;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) eblk[blockIdx.x] = red[0];
}


// reciprocal-space and self terms per site (:2113-2265, :2472-2508)
__global__ void k_recip_site(int n, const SiteParams S, ElecParams P, const double* __restrict__ phi_fix,
                             const double* __restrict__ phi_ind, const double* __restrict__ mu, double* __restrict__ force,
                             double* __restrict__ V, double* __restrict__ eblk) {
    critical_path_wave();
    __shared__ double red[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double en = 0.0;
    if (i < n) {
        const double f = el_electric;
        const double sc[3] = {P.grid.x * P.inv_box.x, P.grid.y * P.inv_box.y, P.grid.z * P.inv_box.z};
        const double* pf = phi_fix + 10 * i;
        const double* pd = phi_ind + 10 * i;
        const double q = S.q[i];
        const double m[3] = {mu[3 * i], mu[3 * i + 1], mu[3 * i + 2]};
        // second-derivative component index [a][k]
        const int d2[3][3] = {{4, 7, 8}, {7, 5, 9}, {8, 9, 6}};
        double fv[3];
        for (int a = 0; a < 3; ++a) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += 2.0 * m[k] * sc[k] * (pf[d2[a][k]] + pd[d2[a][k]]);
            fv[a] = sc[a] * (s + 2.0 * q * pd[1 + a]);
        }
        for (int a = 0; a < 3; ++a) force[3 * i + a] -= fv[a] * 0.5 * f + q * pf[1 + a] * sc[a] * f;
        // energies: induced 0.5 f sum mu.grad(phi_fix) (fractional), fixed 0.5 f q phi, self
        en = 0.5 * f * (sc[0] * m[0] * pf[1] + sc[1] * m[1] * pf[2] + sc[2] * m[2] * pf[3]);
        en += 0.5 * f * q * pf[0];
        en += q * q * -(f * P.alpha / el_sqrt_pi);
        V[i] += pf[0] + pd[0] + q * -(P.alpha / el_sqrt_pi) * 2;
    }
    red[threadIdx.x] = en;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) eblk[blockIdx.x] = red[0];
}

// charge-derivative chain rule (:2767-2773): F_a -= f sum_s dq_s/dr_a V_s, a in O,H1,H2
__global__ void k_charge_deriv_forces(int nw, const double* __restrict__ dq, const double* __restrict__ V,
                                      double* __restrict__ force) {
    critical_path_wave();
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const int base = 4 * w;
    const double Vs[3] = {V[base + 1], V[base + 2], V[base + 3]};
    for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 3; ++d) {
            double s = 0;
            for (int q = 0; q < 3; ++q) s += dq[(size_t)w * 27 + (a * 3 + q) * 3 + d] * Vs[q];
            force[3 * (base + a) + d] += s * -el_electric;
        }
}

// sum block partials into the fixed-point energy slot (fixed order)
__global__ void k_energy_finish(int nblocks, const double* __restrict__ eblk, u64* __restrict__ energy) {
    critical_path_wave();
    __shared__ double red[256];
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += eblk[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(energy, to_fix(red[0]));
}

// double forces of the electrostatics particles -> fixed-point particle forces
__global__ void k_add_forces_fix(int n, const double* __restrict__ f, u64* __restrict__ fx) {
    critical_path_wave();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 3 * n) atomicAdd(&fx[t], to_fix(f[t]));
}

// potential of the charges and induced dipoles at arbitrary points
// (calculateElectrostaticPotentialForParticleGridPoint, :1023-1086): plain
// Coulomb sum over every site, nearest image under PME
__global__ void k_potential_at_points(int npts, const double* __restrict__ pts, int n, const double* __restrict__ pos,
                                      const double* __restrict__ q, const double* __restrict__ mu, ElecParams P,
                                      double* __restrict__ out) {
    critical_path_wave();
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npts) return;
    double v = 0.0;
    for (int i = 0; i < n; ++i) {
        double dx = pos[3 * i] - pts[3 * k], dy = pos[3 * i + 1] - pts[3 * k + 1], dz = pos[3 * i + 2] - pts[3 * k + 2];
        min_image(dx, dy, dz, P);
        const double r2 = dx * dx + dy * dy + dz * dz, r = sqrt(r2);
        const double rr1 = 1.0 / r, rr3 = rr1 / r2;
        v += q[i] * rr1 - (mu[3 * i] * dx + mu[3 * i + 1] * dy + mu[3 * i + 2] * dz) * rr3;
    }
    out[k] = v * el_electric;
}

// System multipole moments about the centre of mass (MBPolReferenceElectrostaticsForce.cpp:923-1021),
// one block of 256 lanes: per-lane partial sums over a fixed stride, then a fixed-order tree
// reduction in LDS, so the result does not depend on scheduling.  out[13]: net charge, dipole
// (Debye), traceless quadrupole (Buckingham, row-major 3x3).
__global__ void __launch_bounds__(256) k_system_moments(int n, const double* __restrict__ pos,
                                                        const double* __restrict__ q, const double* __restrict__ mu,
                                                        const double* __restrict__ mass, double* __restrict__ out) {
    __shared__ double red[10][256];
    const int t = threadIdx.x;
    // centre of mass
    double m = 0, cx = 0, cy = 0, cz = 0;
    for (int i = t; i < n; i += 256) {
        m += mass[i];
        cx += pos[3 * i] * mass[i];
        cy += pos[3 * i + 1] * mass[i];
        cz += pos[3 * i + 2] * mass[i];
    }
    red[0][t] = m; red[1][t] = cx; red[2][t] = cy; red[3][t] = cz;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (t < k)
            for (int c = 0; c < 4; ++c) red[c][t] += red[c][t + k];
        __syncthreads();
    }
    const double mt = red[0][0];
    const double s = mt > 0.0 ? 1.0 / mt : 1.0;
    const double com[3] = {mt > 0.0 ? red[1][0] * s : 0.0, mt > 0.0 ? red[2][0] * s : 0.0, mt > 0.0 ? red[3][0] * s : 0.0};
    __syncthreads();
    // net charge, dipole and second moments
    double a[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = t; i < n; i += 256) {
        const double r[3] = {pos[3 * i] - com[0], pos[3 * i + 1] - com[1], pos[3 * i + 2] - com[2]};
        const double qi = q[i], m0 = mu[3 * i], m1 = mu[3 * i + 1], m2 = mu[3 * i + 2];
        a[0] += qi;
        a[1] += r[0] * qi + m0;
        a[2] += r[1] * qi + m1;
        a[3] += r[2] * qi + m2;
        a[4] += r[0] * r[0] * qi + 2.0 * r[0] * m0;                 // xx
        a[5] += r[0] * r[1] * qi + r[0] * m1 + r[1] * m0;            // xy
        a[6] += r[0] * r[2] * qi + r[0] * m2 + r[2] * m0;            // xz
        a[7] += r[1] * r[1] * qi + 2.0 * r[1] * m1;                 // yy
        a[8] += r[1] * r[2] * qi + r[1] * m2 + r[2] * m1;            // yz
        a[9] += r[2] * r[2] * qi + 2.0 * r[2] * m2;                 // zz
    }
    for (int c = 0; c < 10; ++c) red[c][t] = a[c];
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (t < k)
            for (int c = 0; c < 10; ++c) red[c][t] += red[c][t + k];
        __syncthreads();
    }
    if (t == 0) {
        const double debye = 4.80321;
        const double xx = red[4][0], xy = red[5][0], xz = red[6][0], yy = red[7][0], yz = red[8][0], zz = red[9][0];
        const double tr = (xx + yy + zz) / 3.0, qs = 100.0 * 3.0 * debye;
        out[0] = red[0][0];
        for (int c = 0; c < 3; ++c) out[1 + c] = red[1 + c][0] * 10.0 * debye;
        out[4] = 0.5 * (xx - tr) * qs;
        out[5] = out[7] = 0.5 * xy * qs;
        out[6] = out[10] = 0.5 * xz * qs;
        out[8] = 0.5 * (yy - tr) * qs;
        out[9] = out[11] = 0.5 * yz * qs;
        out[12] = 0.5 * (zz - tr) * qs;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }
constexpr int kG = 16;   // lanes per site: prefactors and fixed field, energy
constexpr int kGI = 32;  // lanes per site in the iteration (>= 25 for the PME gather)

void launch_copy_if_converged(hipStream_t s, long long n, const SolveCtl* ctl, const double* src, double* dst) {
    if (n > 0)
        hipLaunchKernelGGL(k_copy_if_converged, (int)std::min<long long>((n + 255) / 256, 1024), 256, 0, s, n, ctl, src,
                           dst);
}
void launch_site_charges_copy(hipStream_t s, int n, const double* q0, double* q) {
    hipLaunchKernelGGL(k_site_charges_copy, nblk(n, 256), 256, 0, s, n, q0, q);
}
void launch_site_cells(hipStream_t s, int n, const double* pos, ElecParams P, int3 nc, int* cell_of, int* cell_count) {
    hipLaunchKernelGGL(k_site_cells, nblk(n, 256), 256, 0, s, n, pos, P, nc, cell_of, cell_count);
}
void launch_bucket_fill(hipStream_t s, int n, const int* bucket_of, const int* bucket_start, int* bucket_fill,
                        int* members) {
    hipLaunchKernelGGL(k_bucket_fill, nblk(n, 256), 256, 0, s, n, bucket_of, bucket_start, bucket_fill, members);
}
void launch_rank_sort(hipStream_t s, int n, const int* bucket_of, const int* bucket_start, const int* members_in,
                      int* members_out) {
    hipLaunchKernelGGL(k_rank_sort, nblk(n, 256), 256, 0, s, n, bucket_of, bucket_start, members_in, members_out);
}
void launch_site_pack(hipStream_t s, int n, const int* members, const double* pos, const double* pol, double4* cpos) {
    if (n > 0) hipLaunchKernelGGL(k_site_pack, nblk(n, 256), 256, 0, s, n, members, pos, pol, cpos);
}
void launch_site_lists(hipStream_t s, int lanes, int n, const double* pos, const double* pol, ElecParams P, int3 nc,
                       int reach, int use_cells, const int* cell_of, const int* cell_start, const int* members,
                       const double4* cpos, int capP, int capQ, int* nbP, int* cntP, int* nbQ, int* cntQ,
                       int* maxcnt) {
#define MBPOL_SITE_LISTS(G)                                                                                          \
    hipLaunchKernelGGL(k_site_lists<G>, nblk((long long)n * G, 256), 256, 0, s, n, pos, pol, P, nc, reach, use_cells, \
                       cell_of, cell_start, members, cpos, capP, capQ, nbP, cntP, nbQ, cntQ, maxcnt)
    switch (lanes) {
        case 8: MBPOL_SITE_LISTS(8); break;
        case 32: MBPOL_SITE_LISTS(32); break;
        default: MBPOL_SITE_LISTS(16); break;
    }
#undef MBPOL_SITE_LISTS
}
void launch_prefactors_field(hipStream_t s, int npol, const int* polidx, const double* pos, SiteParams S, ElecParams P,
                             const ElecRows& R, double2* pf, double* field) {
    if (npol <= 0) return;
    if (P.pme)
        hipLaunchKernelGGL((k_prefactors_field<kG, true>), nblk((long long)npol * kG, 256), 256, 0, s, npol, polidx,
                           pos, S, P, R.capP, R.nbP, R.cntP, R.capQ, R.nbQ, R.cntQ, pf, field);
    else
        hipLaunchKernelGGL((k_prefactors_field<kG, false>), nblk((long long)npol * kG, 256), 256, 0, s, npol, polidx,
                           pos, S, P, R.capP, R.nbP, R.cntP, R.capQ, R.nbQ, R.cntQ, pf, field);
}
void launch_site_bins(hipStream_t s, int n, const double* pos, const double* weight, ElecParams P, BrickGeom B,
                      int* bin_of, int* bin_count) {
    hipLaunchKernelGGL(k_site_bins, nblk(n, 256), 256, 0, s, n, pos, weight, P, B, bin_of, bin_count);
}
void launch_site_spline_table(hipStream_t s, int n, const double* pos, ElecParams P, int* g0tab, double* sptab) {
    hipLaunchKernelGGL(k_site_spline_table, nblk(n, 128), 128, 0, s, n, pos, P, g0tab, sptab);
}
void launch_spread_bricks(hipStream_t s, int mode, const double* q, const double* mu, ElecParams P, BrickGeom B,
                          const int* bin_start, const int* members, const int* g0tab, const double* sptab,
                          double* grid) {
    hipLaunchKernelGGL(k_spread_bricks, B.nb[0] * B.nb[1] * B.nb[2], kBrickThreads, 0, s, mode, q, mu, P, B, bin_start,
                       members, g0tab, sptab, grid);
}
void launch_csr_count(hipStream_t s, int npol, const int* polidx, const int* g0tab, ElecParams P, int* rowcount) {
    if (npol > 0)
        hipLaunchKernelGGL(k_csr_count, nblk((long long)npol * 125, 256), 256, 0, s, npol, polidx, g0tab, P, rowcount);
}
void launch_csr_fill(hipStream_t s, int npol, const int* polidx, const int* g0tab, const double* sptab, ElecParams P,
                     const int* rowstart, int* fill, int* esite, double* ecoef) {
    if (npol > 0)
        hipLaunchKernelGGL(k_csr_fill, nblk((long long)npol * 125, 256), 256, 0, s, npol, polidx, g0tab, sptab, P,
                           rowstart, fill, esite, ecoef);
}
void launch_csr_sort(hipStream_t s, size_t ngrid, const int* rowstart, int* esite, double* ecoef) {
    hipLaunchKernelGGL(k_csr_sort, nblk(ngrid, 256), 256, 0, s, ngrid, rowstart, esite, ecoef);
}
void launch_spread_csr(hipStream_t s, size_t ngrid, const int* rowstart, const int* esite, const double* ecoef,
                       const double* mu, double* grid, const KEvents& ev, const int* stop) {
    MBPOL_LAUNCH(ev, k_spread_csr, nblk(ngrid, 256), 256, 0, s, ngrid, rowstart, esite, ecoef, mu, grid, stop);
}
void launch_convolve(hipStream_t s, size_t nhalf, const double* influence, double2* spec) {
    hipLaunchKernelGGL(k_convolve, nblk(nhalf, 256), 256, 0, s, nhalf, influence, spec);
}
void launch_influence(hipStream_t s, ElecParams P, int khp, const double* bx, const double* by, const double* bz,
                      double* out) {
    const size_t total = (size_t)P.grid.x * P.grid.y * khp;
    hipLaunchKernelGGL(k_influence, nblk(total, 256), 256, 0, s, P, khp, bx, by, bz, out);
}
void launch_gather(hipStream_t s, int n, const double* pos, ElecParams P, const double* grid, double* phi) {
    hipLaunchKernelGGL(k_gather, nblk(n, 64), 64, 0, s, n, pos, P, grid, phi);
}
void launch_record_field(hipStream_t s, int n, ElecParams P, const double* phi, int accumulate, double self_term,
                         const double* mu, double* field) {
    hipLaunchKernelGGL(k_record_field, nblk(n, 256), 256, 0, s, n, P, phi, accumulate, self_term, mu, field);
}
int iterate_blocks(int npol) { return std::max(1, nblk((long long)npol * kGI, 256)); }
void launch_iterate(hipStream_t s, int npol, const int* polidx, const double* pos, ElecParams P, const ElecRows& R,
                    const double* mu_in, double* mu_out, const double* efix_pol, const double* pol, const double* grid,
                    double self_term, double omega, double* partial, const SolveCtl* ctl, const KEvents& ev) {
    MBPOL_LAUNCH(ev, k_iterate<kGI>, iterate_blocks(npol), 256, 0, s, npol, polidx, pos, P, R.capP, R.nbP, R.cntP, R.pf,
                 mu_in, mu_out, efix_pol, pol, grid, self_term, omega, partial, ctl);
}
int cg_blocks(int nitems) { return std::max(1, nblk(nitems, 256)); }
int precond_wgs(int nblocks) { return cg_blocks(nblocks * kGB); }
int block_ld() { return kBlockLd; }
void launch_block_factor(hipStream_t s, const PrecondBlocks& B, const double* pos, SiteParams S, ElecParams P,
                         double* minv) {
    if (B.n > 0)
        hipLaunchKernelGGL(k_block_factor, precond_wgs(B.n), 256, 0, s, B.n, B.start, B.site, pos, S, P, minv);
}
void launch_cg_matvec(hipStream_t s, int npol, const int* polidx, const double* pos, ElecParams P, const ElecRows& R,
                      const double* v, const double* grid, double self_term, const double* pol, const double* efix,
                      int mode, double* out, double* partial, const SolveCtl* ctl, const KEvents& ev) {
    if (npol > 0)
        MBPOL_LAUNCH(ev, k_cg_matvec<kGI>, iterate_blocks(npol), 256, 0, s, npol, polidx, pos, P, R.capP, R.nbP, R.cntP,
                     R.pf, v, grid, self_term, pol, efix, mode, out, partial, ctl);
}
void launch_cg_precond(hipStream_t s, const PrecondBlocks& B, const double* minv, const double* pol, const double* r,
                       double* z, double* partial, double* blk_ee, const SolveCtl* ctl) {
    if (B.n > 0)
        hipLaunchKernelGGL(k_cg_precond, precond_wgs(B.n), 256, 0, s, B.n, B.start, B.site, minv, pol, r, z, partial,
                           blk_ee, ctl);
}
void launch_cg_update(hipStream_t s, const PrecondBlocks& B, const double* minv, const double* pol, int npol,
                      const double* partial_mv, const double* rz, int k, double* x, double* r, const double* p,
                      const double* q, double* z, double* partial, double* blk_ee, const SolveCtl* ctl) {
    if (B.n > 0)
        hipLaunchKernelGGL(k_cg_update, precond_wgs(B.n), 256, 0, s, B.n, B.start, B.site, minv, pol,
                           iterate_blocks(npol), partial_mv, rz, k, x, r, p, q, z, partial, blk_ee, ctl);
}
void launch_cg_direction(hipStream_t s, int npol, const int* polidx, int nb, const double* partial, double* rz, int k,
                         const double* z, double* p, int n_all, double tol, int max_iter, const GroupConv& gc,
                         SolveCtl* ctl) {
    if (npol > 0)
        hipLaunchKernelGGL(k_cg_direction, cg_blocks(npol), 256, 0, s, npol, polidx, nb, partial, rz, k, z, p, n_all,
                           tol, max_iter, gc, ctl);
}
void launch_sor_finish(hipStream_t s, int nblocks, int n_all, const double* partial, int t, double tol, int max_iter,
                       SolveCtl* ctl) {
    hipLaunchKernelGGL(k_sor_finish, 1, 256, 0, s, nblocks, n_all, partial, t, tol, max_iter, ctl);
}
void launch_scale_copy(hipStream_t s, int n, const double* a, const double* pol, double* out, double* out2) {
    hipLaunchKernelGGL(k_scale_copy, nblk(3 * n, 256), 256, 0, s, 3 * n, a, pol, out, out2);
}
int energy_blocks(int n) { return nblk((long long)n * kG, 256); }
void launch_elec_energy(hipStream_t s, int n, const double* pos, SiteParams S, ElecParams P, const ElecRows& R,
                        const double* mu, double* force, double* V, double* eblk, double* esite) {
    if (P.pme)
        hipLaunchKernelGGL((k_elec_energy<kG, true>), energy_blocks(n), 256, 0, s, n, pos, S, P, R.capP, R.nbP, R.cntP,
                           R.capQ, R.nbQ, R.cntQ, mu, force, V, eblk, esite);
    else
        hipLaunchKernelGGL((k_elec_energy<kG, false>), energy_blocks(n), 256, 0, s, n, pos, S, P, R.capP, R.nbP, R.cntP,
                           R.capQ, R.nbQ, R.cntQ, mu, force, V, eblk, esite);
}
int sor_blocks(int n) { return nblk(n, 256); }
void launch_recip_site(hipStream_t s, int n, SiteParams S, ElecParams P, const double* phi_fix, const double* phi_ind,
                       const double* mu, double* force, double* V, double* eblk) {
    hipLaunchKernelGGL(k_recip_site, nblk(n, 256), 256, 0, s, n, S, P, phi_fix, phi_ind, mu, force, V, eblk);
}
void launch_charge_deriv_forces(hipStream_t s, int nw, const double* dq, const double* V, double* force) {
    hipLaunchKernelGGL(k_charge_deriv_forces, nblk(nw, 64), 64, 0, s, nw, dq, V, force);
}
void launch_energy_finish(hipStream_t s, int nblocks, const double* eblk, u64* energy) {
    hipLaunchKernelGGL(k_energy_finish, 1, 256, 0, s, nblocks, eblk, energy);
}
void launch_potential_at_points(hipStream_t s, int npts, const double* pts, int n, const double* pos, const double* q,
                                const double* mu, ElecParams P, double* out) {
    hipLaunchKernelGGL(k_potential_at_points, nblk(npts, 64), 64, 0, s, npts, pts, n, pos, q, mu, P, out);
}
void launch_system_moments(hipStream_t s, int n, const double* pos, const double* q, const double* mu,
                           const double* mass, double* out) {
    hipLaunchKernelGGL(k_system_moments, 1, 256, 0, s, n, pos, q, mu, mass, out);
}
void launch_add_forces_fix(hipStream_t s, int n, const double* f, u64* fx) {
    hipLaunchKernelGGL(k_add_forces_fix, nblk(3 * n, 256), 256, 0, s, n, f, fx);
}

}  // namespace mbpol
