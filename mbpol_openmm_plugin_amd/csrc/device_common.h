// Device-side helpers shared by the MB-pol HIP kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "kernels.h"

#define MBPOL_TABLE static __constant__
#include "mbpol_params.h"

namespace mbpol {

// Forces and energies are accumulated as 64-bit fixed point (2^32 units per
// kJ/mol/nm or kJ/mol): integer atomics commute, so every evaluation of the
// same input gives bit-identical sums whatever order the waves run in.
constexpr double kFixScale = 4294967296.0;
constexpr double kFixInv = 1.0 / 4294967296.0;

__device__ __forceinline__ unsigned long long to_fix(double v) {
    return (unsigned long long)__double2ll_rn(v * kFixScale);
}

__device__ __forceinline__ void add_force(unsigned long long* f, int atom, double fx, double fy, double fz) {
    atomicAdd(&f[3 * atom + 0], to_fix(fx));
    atomicAdd(&f[3 * atom + 1], to_fix(fy));
    atomicAdd(&f[3 * atom + 2], to_fix(fz));
}

// Sum of N fixed-point values over each run of consecutive lanes with equal
// key (the dimer and trimer lists are ordered by first molecule, so a wave's
// units mostly share a few first molecules).  Integer sums are exact, so one
// atomic per run adds the same bits as one atomic per lane.  Every lane of the
// wave must call it (uniform control flow); returns true on the last lane of
// each run, which then holds the run's sums in v.
template <int N>
__device__ __forceinline__ bool wave_run_sum(long long key, unsigned long long (&v)[N]) {
    const int lane = threadIdx.x & 63;
    const long long prev = __shfl_up(key, 1, 64);
    const unsigned long long heads = __ballot(lane == 0 || prev != key);
    // first lane of this lane's run: the highest run head at or below this lane
    const unsigned long long below = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
    const int start = 63 - __clzll(below);
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const bool take = lane - off >= start;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const unsigned long long o = __shfl_up(v[i], off, 64);
            if (take) v[i] += o;
        }
    }
    return lane == 63 || ((heads >> (lane + 1)) & 1ull);
}

// Sum of N fixed-point values over the lanes of the wave with equal key, in any
// order of the keys: the (key, lane) pairs are sorted across the wave by a
// bitonic network of shuffles (lane breaks ties, so the order is total), each
// lane then takes the values of the lane whose pair it holds, and the runs of
// equal keys are summed as in wave_run_sum.  Every lane must call it; on return
// key holds the sorted key of this lane, and the function returns true on the
// last lane of each run, which holds the run's sums in v.
template <int N>
__device__ __forceinline__ bool wave_sorted_run_sum(long long& key, unsigned long long (&v)[N]) {
    const int lane = threadIdx.x & 63;
    int src = lane;
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const long long pk = __shfl_xor(key, j, 64);
            const int ps = __shfl_xor(src, j, 64);
            const bool partner_less = pk < key || (pk == key && ps < src);
            const bool ascending = (lane & k) == 0, lower = (lane & j) == 0;
            // the lower lane of a pair keeps the smaller pair in an ascending block, the larger in a
            // descending one; the upper lane the other
            if (lower == ascending ? partner_less : !partner_less) {
                key = pk;
                src = ps;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = __shfl(v[i], src, 64);
    return wave_run_sum(key, v);
}

// Wave-level sum (64 lanes).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Energy of each lane to fixed point first, then an integer wave sum and one
// atomic per wave: the total does not depend on which units share a wave.
__device__ __forceinline__ void add_energy(unsigned long long* e, double v) {
    unsigned long long f = to_fix(v);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) f += __shfl_xor(f, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(e, f);
}

// Waves of the electrostatics kernels (the evaluation's critical path) raise
// their issue priority: the short-range kernels run beside them on a second
// stream and would otherwise take turns with them at every SIMD's arbiter.
#ifndef MBPOL_WAVE_PRIO
#define MBPOL_WAVE_PRIO 1
#endif
__device__ __forceinline__ void critical_path_wave() {
#if MBPOL_WAVE_PRIO
    __builtin_amdgcn_s_setprio(2);
#endif
}

__device__ __forceinline__ d3 load_d3(const double* p, int i) { return d3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

// Minimum image of p relative to ref, component-wise: p += L * floor((ref - p)/L + 0.5)
// (platforms/reference/src/MBPolReferenceTwoBodyForce.cpp:66-76).
__device__ __forceinline__ double image_coord(double ref, double p, double L) {
    const double d = ref - p;
    return p + L * floor(d / L + 0.5);
}

__device__ __forceinline__ d3 image_to(const d3& ref, const d3& p, const d3& L) {
    return d3{image_coord(ref.x, p.x, L.x), image_coord(ref.y, p.y, L.y), image_coord(ref.z, p.z, L.z)};
}

template <typename T>
struct v3 {
    T x, y, z;
    __device__ __forceinline__ v3() {}
    __device__ __forceinline__ v3(T a, T b, T c) : x(a), y(b), z(c) {}
    __device__ __forceinline__ v3 operator+(const v3& o) const { return v3(x + o.x, y + o.y, z + o.z); }
    __device__ __forceinline__ v3 operator-(const v3& o) const { return v3(x - o.x, y - o.y, z - o.z); }
    __device__ __forceinline__ v3 operator*(T s) const { return v3(x * s, y * s, z * s); }
    __device__ __forceinline__ v3& operator+=(const v3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    __device__ __forceinline__ v3& operator-=(const v3& o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
    __device__ __forceinline__ T dot(const v3& o) const { return x * o.x + y * o.y + z * o.z; }
    __device__ __forceinline__ v3 cross(const v3& o) const {
        return v3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x);
    }
};

template <typename T>
__device__ __forceinline__ T exp_t(T x);
template <>
__device__ __forceinline__ float exp_t<float>(float x) { return expf(x); }
template <>
__device__ __forceinline__ double exp_t<double>(double x) { return exp(x); }

template <typename T>
__device__ __forceinline__ T rsqrt_t(T x);
template <>
__device__ __forceinline__ float rsqrt_t<float>(float x) { return 1.0f / sqrtf(x); }
template <>
__device__ __forceinline__ double rsqrt_t<double>(double x) { return 1.0 / sqrt(x); }

}  // namespace mbpol
