// PME grid transforms (kernels_fft.hip).
#pragma once
#include "kernels.h"

namespace mbpol {

constexpr int kFftThreads = 128;  // threads per workgroup of the grid transforms

// One axis of the grid: length n = fac[0] * ... * fac[nfac-1] (radices 2, 3,
// 4, 5, 7, 11, 13); stage s has ns[s] = fac[0] ... fac[s-1], nb[s] = n / fac[s]
// butterflies and twiddle step n / (ns[s] fac[s]).  2^lgP threads work on one
// line (lgP <= 7); tw[t] = exp(-2 pi i t / n) in device memory.
struct FftAxis {
    int n;
    int nfac;
    int fac[12], ns[12], nb[12], step[12];
    float inv_ns[12];
    int lgP;
    const double2* tw;
};

struct PmeFft {
    FftAxis ax[3];  // x, y, z
};

// longest axis the LDS transforms take: one line per workgroup, two buffers in 32 KB
constexpr int kFftMaxLength = 1024;
// real grid -> half spectrum -> times influence -> real grid, in place on `real` (H is the spectrum buffer)
void launch_pme_convolution(hipStream_t s, const PmeFft& F, double* real, double2* H, const double* influence);

}  // namespace mbpol
