// Calibration program for hardware memory counters: copies a 256 MiB buffer
// into another with a plain vectorised kernel, three times.  The FETCH_SIZE /
// WRITE_SIZE a profiler reports for one dispatch, divided by 256 MiB, is the
// factor between those counters and HBM bytes.  Not part of the library.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void __launch_bounds__(256) copy_kernel(const double4* __restrict__ in, double4* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

int main() {
    const size_t bytes = 256ull << 20, n = bytes / sizeof(double4);
    double4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        std::fprintf(stderr, "allocation failed\n");
        return 1;
    }
    if (hipMemset(a, 0, bytes) != hipSuccess) return 1;
    for (int r = 0; r < 3; ++r) hipLaunchKernelGGL(copy_kernel, (unsigned)((n + 255) / 256), 256, 0, 0, a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) {
        std::fprintf(stderr, "copy failed\n");
        return 1;
    }
    std::printf("copied %zu bytes per dispatch, 3 dispatches\n", bytes);
    hipFree(a);
    hipFree(b);
    return 0;
}
