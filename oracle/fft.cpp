// Complex FFT for the PME oracle -- test infrastructure only.
//
// The reference uses OpenMM's fftpack (MBPolReferenceElectrostaticsForce.cpp:
// 1420-1422, 2332-2334), unnormalised in both directions.  This is a plain
// recursive mixed-radix Cooley-Tukey transform (smallest prime factor first,
// direct DFT for prime lengths), applied axis by axis.
#include <complex>
#include <vector>
#include "oracle_common.h"

namespace oracle {
typedef std::complex<double> cplx;

static void fft1d(cplx* data, int n, int stride, int sign, std::vector<cplx>& scratch) {
    if (n == 1) return;
    int p = 2;
    while (p * p <= n && n % p) ++p;
    if (n % p) p = n;  // prime length
    const int m = n / p;
    std::vector<cplx> tmp(n);
    for (int i = 0; i < n; ++i) tmp[i] = data[(size_t)i * stride];
    if (p == n) {
        for (int k = 0; k < n; ++k) {
            cplx s = 0;
            for (int j = 0; j < n; ++j) {
                double ang = sign * 2.0 * M_PI * (double)((long long)j * k % n) / n;
                s += tmp[j] * cplx(std::cos(ang), std::sin(ang));
            }
            data[(size_t)k * stride] = s;
        }
        return;
    }
    // split into p interleaved sub-sequences of length m
    std::vector<cplx> sub((size_t)p * m);
    for (int r = 0; r < p; ++r)
        for (int j = 0; j < m; ++j) sub[(size_t)r * m + j] = tmp[(size_t)j * p + r];
    for (int r = 0; r < p; ++r) fft1d(&sub[(size_t)r * m], m, 1, sign, scratch);
    for (int k = 0; k < n; ++k) {
        cplx s = 0;
        const int km = k % m;
        for (int r = 0; r < p; ++r) {
            double ang = sign * 2.0 * M_PI * (double)((long long)r * k % n) / n;
            s += sub[(size_t)r * m + km] * cplx(std::cos(ang), std::sin(ang));
        }
        data[(size_t)k * stride] = s;
    }
}

// In-place 3D transform of an (nx, ny, nz) row-major grid; sign -1 forward.
void fft3d(cplx* grid, int nx, int ny, int nz, int sign) {
    std::vector<cplx> scratch;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y) fft1d(grid + ((size_t)x * ny + y) * nz, nz, 1, sign, scratch);
    for (int x = 0; x < nx; ++x)
        for (int z = 0; z < nz; ++z) fft1d(grid + (size_t)x * ny * nz + z, ny, nz, sign, scratch);
    for (int y = 0; y < ny; ++y)
        for (int z = 0; z < nz; ++z) fft1d(grid + (size_t)y * nz + z, nx, ny * nz, sign, scratch);
}

}  // namespace oracle
