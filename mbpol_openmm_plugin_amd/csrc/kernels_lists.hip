// O-O cell lists and the stream-compacted dimer / chain-trimer lists.
//
// Set semantics follow the reference exactly, with FP64 distances computed in
// the same operation order (no contraction) so that boundary cases fall the
// same way:
//  * dimers: every pair i<j with d^2 <= rc^2 (minimum image when periodic) --
//    OpenMM's computeNeighborListVoxelHash as called at
//    platforms/reference/src/MBPolReferenceKernels.cpp:379-383;
//  * trimers: the "chain" rule of ReferenceThreeNeighborList.cpp:181-226 --
//    (i, j, k) with j in lo(i), k in lo(j), lo(i) = {j < i : d(i,j) <= rc},
//    d(i,k) never tested; a molecule whose wrapped coordinate rounds into the
//    voxel index nv is invisible to later molecules there (:86-97, :140-146),
//    which is reproduced.
// Lists are built in two passes (count, exclusive scan, fill); the fill pass
// recomputes the same predicate, so it writes exactly `count` entries -- those
// below the list's capacity: the host learns the totals after the evaluation
// and runs it again with larger lists when one did not fit, so that no host
// wait sits between the passes.
#include "device_common.h"
#include "kernels.h"

namespace mbpol {

#pragma clang fp contract(off)

// periodicDifference (ReferenceThreeNeighborList.cpp:14-18)
__device__ __forceinline__ double pdiff(double a, double b, double L) {
    const double diff = a - b;
    const double base = floor(diff / L + 0.5) * L;
    return diff - base;
}

// squared distance p1 -> p2 (ReferenceThreeNeighborList.cpp:21-34)
__device__ __forceinline__ double dist2(const d3& p1, const d3& p2, const d3& L, int periodic) {
    double dx, dy, dz;
    if (!periodic) {
        dx = p2.x - p1.x;
        dy = p2.y - p1.y;
        dz = p2.z - p1.z;
    } else {
        dx = pdiff(p2.x, p1.x, L.x);
        dy = pdiff(p2.y, p1.y, L.y);
        dz = pdiff(p2.z, p1.z, L.z);
    }
    return dx * dx + dy * dy + dz * dz;
}

__device__ __forceinline__ d3 opos(const double* pos, const int* idx, int m) { return load_d3(pos, idx[3 * m]); }

__device__ __forceinline__ int wrap_cell(int c, int n) { return c < 0 ? c + n : (c >= n ? c - n : c); }

// ---- cell binning (periodic, nc >= 3 per dimension) ----
__global__ void k_cell_count(int n, const int* __restrict__ idx, const double* __restrict__ pos, d3 L, int3 nc,
                             int* __restrict__ cell_of, int* __restrict__ cell_count) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const d3 p = opos(pos, idx, m);
    const double q[3] = {p.x - L.x * floor(p.x / L.x), p.y - L.y * floor(p.y / L.y), p.z - L.z * floor(p.z / L.z)};
    const double Ld[3] = {L.x, L.y, L.z};
    const int ncd[3] = {nc.x, nc.y, nc.z};
    int c[3];
    for (int d = 0; d < 3; ++d) {
        int v = (int)floor(q[d] * ncd[d] / Ld[d]);
        c[d] = v < 0 ? 0 : (v >= ncd[d] ? ncd[d] - 1 : v);
    }
    const int cell = (c[0] * nc.y + c[1]) * nc.z + c[2];
    cell_of[m] = cell;
    atomicAdd(&cell_count[cell], 1);
}

__global__ void k_cell_fill(int n, const int* __restrict__ cell_of, const int* __restrict__ cell_start,
                            int* __restrict__ cell_fill, int* __restrict__ cell_members) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const int c = cell_of[m];
    const int slot = atomicAdd(&cell_fill[c], 1);
    cell_members[cell_start[c] + slot] = m;
}

// Lanes of one G-lane group share molecule m; these helpers give the group's
// lane bits and the bits of the lanes below this one, for ballot compaction.
template <int G>
__device__ __forceinline__ unsigned long long group_bits() {
    const int lane = threadIdx.x & 63;
    return G == 64 ? ~0ull : (((1ull << G) - 1ull) << (lane & ~(G - 1)));
}
__device__ __forceinline__ unsigned long long lanes_below() { return (1ull << (threadIdx.x & 63)) - 1ull; }

// Visit every candidate partner of molecule m with the G lanes of its group:
// the 27 neighbouring cells when a cell grid is in use, else every molecule.
// f(j, active) is called by all lanes of the group in step (active = lane has
// a candidate), so that it may use wavefront ballots.
template <int G, typename F>
__device__ __forceinline__ void for_candidates(int m, int lane, int n, const ListGeom& g, const int* cell_of,
                                               const int* cell_start, const int* cell_members, F&& f) {
    if (!g.use_cells) {
        const int b0 = g.gs ? g.gs[m] : 0, b1 = g.gs ? g.ge[m] : n;
        for (int b = b0; b < b1; b += G) f(b + lane, b + lane < b1);
        return;
    }
    // the three cells along z of each (X, Y) column are consecutive in the cell order:
    // one or two contiguous member ranges, visited in the same order as cell by cell
    const int c = cell_of[m];
    const int cz = c % g.nc.z, cy = (c / g.nc.z) % g.nc.y, cx = c / (g.nc.z * g.nc.y);
    auto scan = [&](int base, int za, int zb) {
        const int b = cell_start[base + za], e = cell_start[base + zb + 1];
        for (int t = b; t < e; t += G) {
            const bool act = t + lane < e;
            f(act ? cell_members[t + lane] : -1, act);
        }
    };
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
            const int base = (wrap_cell(cx + dx, g.nc.x) * g.nc.y + wrap_cell(cy + dy, g.nc.y)) * g.nc.z;
            const int z0 = cz - 1, z1 = cz + 1;
            if (z0 < 0) {
                scan(base, z0 + g.nc.z, g.nc.z - 1);
                scan(base, 0, z1);
            } else if (z1 >= g.nc.z) {
                scan(base, z0, g.nc.z - 1);
                scan(base, 0, z1 - g.nc.z);
            } else {
                scan(base, z0, z1);
            }
        }
}

constexpr int kGL = 16;  // lanes per molecule in the pair passes

// ---- dimer list: count pass (pairs == nullptr) and fill pass ----
__global__ void __launch_bounds__(256) k_dimer_pass(int n, const int* __restrict__ idx, const double* __restrict__ pos,
                                                    ListGeom g, const int* __restrict__ cell_of,
                                                    const int* __restrict__ cell_start,
                                                    const int* __restrict__ cell_members, long long* __restrict__ count,
                                                    const long long* __restrict__ offset, int2* __restrict__ pairs,
                                                    long long cap) {
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / kGL, lane = threadIdx.x % kGL;
    if (i >= n) return;  // uniform over the group
    const d3 pi = opos(pos, idx, i);
    const double rc2 = g.rc * g.rc;
    const unsigned long long gb = group_bits<kGL>(), below = lanes_below();
    long long c = 0;
    const long long o = offset ? offset[i] : 0;
    for_candidates<kGL>(i, lane, n, g, cell_of, cell_start, cell_members, [&](int j, bool act) {
        const bool ok = act && j > i && dist2(pi, opos(pos, idx, j), g.L, g.periodic) <= rc2;
        const unsigned long long bits = __ballot(ok) & gb;
        if (ok && pairs) {
            const long long t = o + c + __popcll(bits & below);
            if (t < cap) pairs[t] = make_int2(i, j);  // beyond the capacity: counted, reported, not written
        }
        c += __popcll(bits);
    });
    if (count && lane == 0) count[i] = c;
}

// ---- chain trimer list ----
// voxel index of the reference's ThreeVoxelHash (ReferenceThreeNeighborList.cpp:86-97); a molecule whose
// index equals nv in some dimension is never found from a later molecule
__device__ __forceinline__ bool phantom_voxel(const d3& p, const ListGeom& g) {
    if (!g.periodic) return false;
    const double Ld[3] = {g.L.x, g.L.y, g.L.z}, pd[3] = {p.x, p.y, p.z};
    for (int d = 0; d < 3; ++d) {
        const double edge = Ld[d] / floor(Ld[d] / g.rc);
        const int nv = (int)floor(Ld[d] / edge + 0.5);
        const double q = pd[d] - Ld[d] * floor(pd[d] / Ld[d]);
        if ((int)floor(q / edge) >= nv) return true;
    }
    return false;
}

// lo(i) = {j < i : d(i,j) <= rc, j not in a phantom voxel}: count pass (lower == nullptr) and fill pass
__global__ void __launch_bounds__(256) k_lower_pass(int n, const int* __restrict__ idx, const double* __restrict__ pos,
                                                    ListGeom g, const int* __restrict__ cell_of,
                                                    const int* __restrict__ cell_start,
                                                    const int* __restrict__ cell_members, int* __restrict__ count,
                                                    const int* __restrict__ offset, int* __restrict__ lower, int cap) {
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / kGL, lane = threadIdx.x % kGL;
    if (i >= n) return;
    const d3 pi = opos(pos, idx, i);
    const double rc2 = g.rc * g.rc;
    const unsigned long long gb = group_bits<kGL>(), below = lanes_below();
    int c = 0;
    const int o = offset ? offset[i] : 0;
    for_candidates<kGL>(i, lane, n, g, cell_of, cell_start, cell_members, [&](int j, bool act) {
        bool ok = act && j < i;
        if (ok) {
            const d3 pj = opos(pos, idx, j);
            ok = dist2(pi, pj, g.L, g.periodic) <= rc2 && !phantom_voxel(pj, g);
        }
        const unsigned long long bits = __ballot(ok) & gb;
        if (ok && lower) {
            const int t = o + c + __popcll(bits & below);
            if (t < cap) lower[t] = j;
        }
        c += __popcll(bits);
    });
    if (count && lane == 0) count[i] = c;
}

// O-O distances of a triplet after the reference's imaging (every O imaged to Oa,
// MBPolReferenceTwoBodyForce.cpp:78-109), in Angstrom.
__device__ __forceinline__ void trimer_oo(const d3& a, d3 b, d3 c, const ListGeom& g, double& rab, double& rac, double& rbc) {
    if (g.periodic) {
        b = image_to(a, b, g.L);
        c = image_to(a, c, g.L);
    }
    auto d = [](const d3& u, const d3& v) {
        const double x = (u.x - v.x) * 10., y = (u.y - v.y) * 10., z = (u.z - v.z) * 10.;
        return sqrt(x * x + y * y + z * z);
    };
    rab = d(a, b);
    rac = d(a, c);
    rbc = d(b, c);
}

// A triplet contributes nothing when an O-O distance is below 2 A
// (MBPolReferenceThreeBodyForce.cpp:165-166) or when fewer than two of the
// three switches are non-zero (the switch is 0 for r >= 4.5 A, :103-120).
__device__ __forceinline__ bool trimer_active(const d3& a, const d3& b, const d3& c, const ListGeom& g) {
    double rab, rac, rbc;
    trimer_oo(a, b, c, g, rab, rac, rbc);
    if (rab < 2 || rac < 2 || rbc < 2) return false;
    const double rf = mbpol_params::th_r3f;
    const int on = (rab < rf) + (rac < rf) + (rbc < rf);
    return on >= 2;
}

constexpr int kGT = 8;  // lanes per molecule in the trimer pass

// trimers (i, j, k) with j in lo(i), k in lo(j): the lanes of i's group take
// the k of one j at a time, compacting the active ones by ballot
__global__ void __launch_bounds__(256) k_trimer_pass(int n, const int* __restrict__ idx, const double* __restrict__ pos,
                                                     ListGeom g, const int* __restrict__ lo_start,
                                                     const int* __restrict__ lower, int full_list,
                                                     long long* __restrict__ count, const long long* __restrict__ offset,
                                                     int* __restrict__ trip, int lower_cap, long long cap) {
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / kGT, lane = threadIdx.x % kGT;
    if (i >= n) return;  // uniform over the group
    // an incomplete lower list (it overflowed its capacity): nothing to read; the
    // evaluation is run again with a larger one
    if (lo_start[n] > lower_cap) {
        if (count && lane == 0) count[i] = 0;
        return;
    }
    const d3 pi = opos(pos, idx, i);
    const unsigned long long gb = group_bits<kGT>(), below = lanes_below();
    long long c = 0;
    const long long o = offset ? offset[i] : 0;
    for (int a = lo_start[i]; a < lo_start[i + 1]; ++a) {
        const int j = lower[a];
        const d3 pj = opos(pos, idx, j);
        const int e = lo_start[j + 1];
        for (int b = lo_start[j]; b < e; b += kGT) {
            const bool act = b + lane < e;
            const int k = act ? lower[b + lane] : 0;
            const bool ok = act && (full_list || trimer_active(pi, pj, opos(pos, idx, k), g));
            const unsigned long long bits = __ballot(ok) & gb;
            if (ok && trip) {
                const long long t = o + c + __popcll(bits & below);
                if (t < cap) {
                    trip[3 * t] = i;
                    trip[3 * t + 1] = j;
                    trip[3 * t + 2] = k;
                }
            }
            c += __popcll(bits);
        }
    }
    if (count && lane == 0) count[i] = c;
}

}  // namespace mbpol

namespace mbpol {
static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }
void launch_cell_count(hipStream_t s, int n, const int* idx, const double* pos, d3 L, int3 nc, int* cell_of,
                       int* cell_count) {
    if (n > 0) hipLaunchKernelGGL(k_cell_count, nblk(n, 256), 256, 0, s, n, idx, pos, L, nc, cell_of, cell_count);
}
void launch_cell_fill(hipStream_t s, int n, const int* cell_of, const int* cell_start, int* cell_fill,
                      int* cell_members) {
    if (n > 0) hipLaunchKernelGGL(k_cell_fill, nblk(n, 256), 256, 0, s, n, cell_of, cell_start, cell_fill, cell_members);
}
void launch_dimer_pass(hipStream_t s, int n, const int* idx, const double* pos, ListGeom g, const int* cell_of,
                       const int* cell_start, const int* cell_members, long long* count, const long long* offset,
                       int2* pairs, long long cap) {
    if (n > 0)
        hipLaunchKernelGGL(k_dimer_pass, nblk((long long)n * kGL, 256), 256, 0, s, n, idx, pos, g, cell_of, cell_start,
                           cell_members, count, offset, pairs, cap);
}
void launch_lower_pass(hipStream_t s, int n, const int* idx, const double* pos, ListGeom g, const int* cell_of,
                       const int* cell_start, const int* cell_members, int* count, const int* offset, int* lower,
                       int cap) {
    if (n > 0)
        hipLaunchKernelGGL(k_lower_pass, nblk((long long)n * kGL, 256), 256, 0, s, n, idx, pos, g, cell_of, cell_start,
                           cell_members, count, offset, lower, cap);
}
void launch_trimer_pass(hipStream_t s, int n, const int* idx, const double* pos, ListGeom g, const int* lo_start,
                        const int* lower, int full_list, long long* count, const long long* offset, int* trip,
                        int lower_cap, long long cap) {
    if (n > 0)
        hipLaunchKernelGGL(k_trimer_pass, nblk((long long)n * kGT, 256), 256, 0, s, n, idx, pos, g, lo_start, lower,
                           full_list, count, offset, trip, lower_cap, cap);
}
}  // namespace mbpol
