// Dimer and "chain" trimer neighbour lists -- oracle (test infrastructure only).
#include <algorithm>
#include <map>
#include <tuple>
#include "oracle_common.h"

namespace oracle {

// periodicDifference (platforms/reference/src/ReferenceThreeNeighborList.cpp:14-18).
static double periodic_diff(double a, double b, double period) {
    double diff = a - b;
    double base = std::floor(diff / period + 0.5) * period;
    return diff - base;
}

// Squared distance, minimum image when periodic (ReferenceThreeNeighborList.cpp:21-34).
double pair_dist2(const V3& p1, const V3& p2, const V3& box, bool periodic) {
    double dx, dy, dz;
    if (!periodic) {
        dx = p2[0] - p1[0];
        dy = p2[1] - p1[1];
        dz = p2[2] - p1[2];
    } else {
        dx = periodic_diff(p2[0], p1[0], box[0]);
        dy = periodic_diff(p2[1], p1[1], box[1]);
        dz = periodic_diff(p2[2], p1[2], box[2]);
    }
    return dx * dx + dy * dy + dz * dz;
}

// Dimer list: every unordered O-O pair (i<j) with d^2 <= rc^2.  The reference
// gets it from OpenMM's computeNeighborListVoxelHash (called at
// MBPolReferenceKernels.cpp:379-383) whose pair set is "distance <= cutoff,
// minimum image when periodic"; here it is enumerated by brute force.
int64_t dimer_list(int n, const double* opos, const double* box, bool periodic, double rc, int* out,
                   int64_t max_out) {
    const V3 b = box ? V3(box[0], box[1], box[2]) : V3();
    const double rc2 = rc * rc;
    int64_t count = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            if (pair_dist2(load3(opos, i), load3(opos, j), b, periodic) > rc2) continue;
            if (count < max_out) { out[2 * count] = i; out[2 * count + 1] = j; }
            ++count;
        }
    return count;
}

// Chain trimer list, restating ReferenceThreeNeighborList.cpp:36-226:
// voxel hash with edge rc (non-periodic) or L/floor(L/rc) (periodic); atoms
// are visited in index order and matched against the atoms inserted before
// them, so nbr[i] holds j < i with d(i,j) <= rc; triplets are (i, j, k) with
// j in nbr[i] and k in nbr[j] -- d(i,k) is never tested.
int64_t trimer_list(int n, const double* opos, const double* box, bool periodic, double rc, double rmin,
                    int* out, int64_t max_out) {
    const V3 b = box ? V3(box[0], box[1], box[2]) : V3();
    double edge[3];
    int nv[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) {
        edge[d] = periodic ? b[d] / std::floor(b[d] / rc) : rc;
        if (periodic) nv[d] = (int)std::floor(b[d] / edge[d] + 0.5);
    }
    auto voxel_of = [&](const V3& p) {
        int v[3];
        for (int d = 0; d < 3; ++d) {
            double q = periodic ? p[d] - b[d] * std::floor(p[d] / b[d]) : p[d];
            v[d] = (int)std::floor(q / edge[d]);
        }
        return std::make_tuple(v[0], v[1], v[2]);
    };
    std::map<std::tuple<int, int, int>, std::vector<int>> hash;
    std::vector<std::vector<int>> nbr(n);
    const double max2 = rc * rc, min2 = rmin * rmin;
    for (int a = 0; a < n; ++a) {
        const V3 pa = load3(opos, a);
        auto c = voxel_of(pa);
        int reach[3], last[3], first[3];
        for (int d = 0; d < 3; ++d) {
            reach[d] = (int)(rc / edge[d]) + 1;
            int cd = d == 0 ? std::get<0>(c) : (d == 1 ? std::get<1>(c) : std::get<2>(c));
            first[d] = cd - reach[d];
            last[d] = cd + reach[d];
            if (periodic) last[d] = std::min(last[d], cd - reach[d] + nv[d] - 1);
        }
        for (int x = first[0]; x <= last[0]; ++x)
            for (int y = first[1]; y <= last[1]; ++y)
                for (int z = first[2]; z <= last[2]; ++z) {
                    auto key = periodic ? std::make_tuple((x + nv[0]) % nv[0], (y + nv[1]) % nv[1], (z + nv[2]) % nv[2])
                                        : std::make_tuple(x, y, z);
                    auto it = hash.find(key);
                    if (it == hash.end()) continue;
                    for (int j : it->second) {
                        if (j == a) continue;
                        double d2 = pair_dist2(pa, load3(opos, j), b, periodic);
                        if (d2 > max2) continue;
                        if (d2 < min2) continue;
                        nbr[a].push_back(j);
                    }
                }
        hash[c].push_back(a);
    }
    int64_t count = 0;
    for (int i = 0; i < n; ++i)
        for (int j : nbr[i])
            for (int k : nbr[j]) {
                if (count < max_out) {
                    out[3 * count] = i;
                    out[3 * count + 1] = j;
                    out[3 * count + 2] = k;
                }
                ++count;
            }
    return count;
}

}  // namespace oracle
