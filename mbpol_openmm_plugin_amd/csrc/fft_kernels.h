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
    int inplace;  // every stage has at most 2^lgP butterflies: stages run in place (one LDS buffer)
    const double2* tw;
};

struct PmeFft {
    FftAxis ax[3];  // x, y, z
    int khp;        // padded kz row length of the half spectrum (>= n_z / 2 + 1, multiple of 8)
};

// longest axis the LDS transforms take: one line per workgroup, two buffers in 32 KB
constexpr int kFftMaxLength = 1024;
// real grid -> half spectrum H -> (x pass) times influence -> real grid, as five passes:
// z_r2c, y (forward), x_conv, y (inverse), z_c2r
// stop: when non-null and *stop != 0 at launch, the pass does nothing (a dipole solve that has converged)
void launch_fft_z_r2c(hipStream_t s, const PmeFft& F, const double* real, double2* H, const KEvents& ev,
                      const int* stop = nullptr);
// the dipole spreading (rows of the spreading table in grid-point order: rowstart, esite,
// ecoef; dipoles mu) fused with the z pass: the half spectrum of the dipole grid
void launch_spread_fft_z_r2c(hipStream_t s, const PmeFft& F, const int* rowstart, const int* esite,
                             const double* ecoef, const double* mu, double2* H, const KEvents& ev,
                             const int* stop = nullptr);
void launch_fft_y(hipStream_t s, const PmeFft& F, double2* H, int inverse, const KEvents& ev,
                  const int* stop = nullptr);
void launch_fft_x_conv(hipStream_t s, const PmeFft& F, double2* H, const double* influence, const KEvents& ev,
                       const int* stop = nullptr);
void launch_fft_z_c2r(hipStream_t s, const PmeFft& F, const double2* H, double* real, const KEvents& ev,
                      const int* stop = nullptr);
// z and y passes of one direction as one kernel per x plane (plane in LDS);
// fft_plane_ok: the plane fits in LDS and every stage has at most one butterfly per thread
bool fft_plane_ok(const PmeFft& F);
void launch_fft_plane_fwd(hipStream_t s, const PmeFft& F, const double* real, double2* H, const KEvents& ev,
                          const int* stop = nullptr);
void launch_fft_plane_inv(hipStream_t s, const PmeFft& F, const double2* H, double* real, const KEvents& ev,
                          const int* stop = nullptr);

}  // namespace mbpol
