// Exclusive prefix sums of int and long long arrays (list offsets, cell
// starts, the PME spreading table's row starts), in three kernels: per-block
// totals, a one-block scan of the totals, and the per-block scan with its
// block's offset.  Integer sums, so the result does not depend on scheduling;
// kernels only (no library call and no memset), so an evaluation that uses
// them can be captured in a HIP graph.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace mbpol {
namespace {

constexpr int kThreads = 256, kItems = 4, kTile = kThreads * kItems;

// exclusive scan of one value per thread over a 256-thread block; returns the block total in *total
template <typename T>
__device__ __forceinline__ T block_exclusive(T v, T* lds, T* total) {
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
        const T add = t >= off ? lds[t - off] : T(0);
        __syncthreads();
        lds[t] += add;
        __syncthreads();
    }
    const T incl = lds[t];
    if (total) *total = lds[kThreads - 1];
    __syncthreads();
    return incl - v;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_tile_totals(const T* __restrict__ in, long long n, T* __restrict__ part) {
    __shared__ T lds[kThreads];
    const long long base = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kItems;
    T s = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (base + k < n) s += in[base + k];
    T total;
    block_exclusive(s, lds, &total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// one block: exclusive scan of the tile totals, in place
template <typename T>
__global__ void __launch_bounds__(kThreads) k_scan_totals(T* __restrict__ part, int nt) {
    __shared__ T lds[kThreads];
    const int per = (nt + kThreads - 1) / kThreads;
    const int b = threadIdx.x * per;
    T s = 0;
    for (int k = 0; k < per; ++k)
        if (b + k < nt) s += part[b + k];
    T run = block_exclusive(s, lds, (T*)nullptr);
    for (int k = 0; k < per; ++k)
        if (b + k < nt) {
            const T v = part[b + k];
            part[b + k] = run;
            run += v;
        }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_tile_scan(const T* __restrict__ in, long long n,
                                                        const T* __restrict__ part, T* __restrict__ out) {
    __shared__ T lds[kThreads];
    const long long base = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kItems;
    T v[kItems];
    T s = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        v[k] = base + k < n ? in[base + k] : T(0);
        s += v[k];
    }
    T run = block_exclusive(s, lds, (T*)nullptr) + part[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (base + k < n) {
            out[base + k] = run;
            run += v[k];
        }
}

template <typename T>
void scan(hipStream_t s, const T* in, T* out, long long n, T* part) {
    if (n <= 0) return;
    const int nt = (int)((n + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_tile_totals<T>, nt, kThreads, 0, s, in, n, part);
    hipLaunchKernelGGL(k_scan_totals<T>, 1, kThreads, 0, s, part, nt);
    hipLaunchKernelGGL(k_tile_scan<T>, nt, kThreads, 0, s, in, n, part, out);
}

}  // namespace

size_t scan_scratch_bytes(long long n) { return sizeof(long long) * (size_t)((n + kTile - 1) / kTile + 1); }
void launch_scan_i32(hipStream_t s, const int* in, int* out, long long n, void* scratch) {
    scan<int>(s, in, out, n, static_cast<int*>(scratch));
}
void launch_scan_i64(hipStream_t s, const long long* in, long long* out, long long n, void* scratch) {
    scan<long long>(s, in, out, n, static_cast<long long*>(scratch));
}

}  // namespace mbpol
