// Reciprocal-space convolution of the PME grid: forward real-to-complex 3D DFT,
// multiplication by the influence function, inverse complex-to-real 3D DFT
// (the role hipfftExecD2Z / convolve / hipfftExecZ2D play otherwise; the
// reference does the same with its own complex FFT, MBPolReferenceElectrostaticsForce.cpp:1622-1713).
//
// Five passes over the grid, each a batch of 1D transforms done in LDS by one
// workgroup on a few lines at a time (mixed-radix Stockham, radices 2-13):
//   z: real lines -> half spectrum (kz < K2/2 + 1)
//   y: forward
//   x: forward, times the influence function, inverse (one pass: the x lines
//      are complete in k-space after the forward transform)
//   y: inverse
//   z: half spectrum -> real lines (Hermitian extension)
// The half spectrum is laid out [kx][ky][kz] as by a D2Z transform, with each
// kz row padded to khp entries (a multiple of 8, i.e. 128 bytes) so that the
// kz tiles of the y and x passes fall on whole cache lines; both directions
// are unnormalised.
#include <algorithm>

#include "device_common.h"
#include "fft_kernels.h"

namespace mbpol {
namespace {

// cos / sin of 2 pi t / R, t < R
template <int R>
struct Roots;
template <>
struct Roots<2> {
    static constexpr double c[2] = {1.0, -1.0};
    static constexpr double s[2] = {0.0, 0.0};
};
template <>
struct Roots<3> {
    static constexpr double c[3] = {1.0, -0.5, -0.5};
    static constexpr double s[3] = {0.0, 0.8660254037844386, -0.8660254037844386};
};
template <>
struct Roots<4> {
    static constexpr double c[4] = {1.0, 0.0, -1.0, 0.0};
    static constexpr double s[4] = {0.0, 1.0, 0.0, -1.0};
};
template <>
struct Roots<5> {
    static constexpr double c[5] = {1.0, 0.30901699437494745, -0.8090169943749475, -0.8090169943749475, 0.30901699437494745};
    static constexpr double s[5] = {0.0, 0.9510565162951535, 0.5877852522924731, -0.5877852522924731, -0.9510565162951535};
};
template <>
struct Roots<7> {
    static constexpr double c[7] = {1.0, 0.6234898018587335, -0.2225209339563144, -0.9009688679024191, -0.9009688679024191, -0.2225209339563144, 0.6234898018587335};
    static constexpr double s[7] = {0.0, 0.7818314824680298, 0.9749279121818236, 0.4338837391175581, -0.4338837391175581, -0.9749279121818236, -0.7818314824680298};
};
template <>
struct Roots<11> {
    static constexpr double c[11] = {1.0, 0.8412535328311812, 0.41541501300188644, -0.14231483827328514, -0.6548607339452851, -0.9594929736144974, -0.9594929736144974, -0.6548607339452851, -0.14231483827328514, 0.41541501300188644, 0.8412535328311812};
    static constexpr double s[11] = {0.0, 0.5406408174555976, 0.9096319953545183, 0.9898214418809327, 0.7557495743542583, 0.28173255684142967, -0.28173255684142967, -0.7557495743542583, -0.9898214418809327, -0.9096319953545183, -0.5406408174555976};
};
template <>
struct Roots<13> {
    static constexpr double c[13] = {1.0, 0.8854560256532099, 0.5680647467311558, 0.12053668025532305, -0.3546048870425356, -0.7485107481711011, -0.970941817426052, -0.970941817426052, -0.7485107481711011, -0.3546048870425356, 0.12053668025532305, 0.5680647467311558, 0.8854560256532099};
    static constexpr double s[13] = {0.0, 0.46472317204376856, 0.8229838658936564, 0.992708874098054, 0.9350162426854148, 0.6631226582407952, 0.23931566428755777, -0.23931566428755777, -0.6631226582407952, -0.9350162426854148, -0.992708874098054, -0.8229838658936564, -0.46472317204376856};
};

// One radix-R Stockham step (stage with Ns = product of the earlier radices)
// for butterfly j of a line of length N: inputs at j + r N/R, twiddle
// exp(-+2 pi i (j mod Ns) r / (Ns R)) = tw[(j mod Ns) r step], step = N/(Ns R),
// R-point DFT, outputs at (j / Ns) Ns R + j mod Ns + q Ns.  tw[t] = exp(-2 pi i t / N).
template <int R>
__device__ __forceinline__ void butterfly(const double2* __restrict__ in, double2* __restrict__ out, int j, int jq,
                                          int Ns, int stride, int step, const double2* __restrict__ tw, bool inv) {
    // components kept in plain double arrays (struct copies of double2 would go through scratch memory)
    double vx[R], vy[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        vx[r] = in[j + r * stride].x;
        vy[r] = in[j + r * stride].y;
    }
    const int k = j - jq * Ns;
    if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
            const double wx = tw[k * r * step].x, wy0 = tw[k * r * step].y;
            const double wy = inv ? -wy0 : wy0;
            const double x = vx[r] * wx - vy[r] * wy, y = vx[r] * wy + vy[r] * wx;
            vx[r] = x;
            vy[r] = y;
        }
    }
    const int base = jq * Ns * R + k;
    if constexpr (R % 2 == 0) {  // radices 2 and 4: plain R-point DFT
#pragma unroll
        for (int q = 0; q < R; ++q) {
            double ax = vx[0], ay = vy[0];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const int t = (r * q) % R;
                const double c = Roots<R>::c[t], sn = inv ? Roots<R>::s[t] : -Roots<R>::s[t];
                ax += vx[r] * c - vy[r] * sn;
                ay += vx[r] * sn + vy[r] * c;
            }
            out[base + q * Ns].x = ax;
            out[base + q * Ns].y = ay;
        }
    } else {
        // odd R: with s_r = v_r + v_(R-r), d_r = v_r - v_(R-r) (r = 1..(R-1)/2),
        // X_q = v_0 + sum_r s_r cos(2 pi r q / R) -+ i sum_r d_r sin(2 pi r q / R)
        // and X_(R-q) the same with the sign of the sine term flipped
        constexpr int H = (R - 1) / 2;
        double sx[H], sy[H], dx[H], dy[H];
        double x0 = vx[0], y0 = vy[0];
#pragma unroll
        for (int r = 1; r <= H; ++r) {
            sx[r - 1] = vx[r] + vx[R - r];
            sy[r - 1] = vy[r] + vy[R - r];
            dx[r - 1] = vx[r] - vx[R - r];
            dy[r - 1] = vy[r] - vy[R - r];
            x0 += sx[r - 1];
            y0 += sy[r - 1];
        }
        out[base].x = x0;
        out[base].y = y0;
#pragma unroll
        for (int q = 1; q <= H; ++q) {
            double ax = vx[0], ay = vy[0], bx = 0.0, by = 0.0;
#pragma unroll
            for (int r = 1; r <= H; ++r) {
                const int t = (r * q) % R;
                const double c = Roots<R>::c[t], sn = Roots<R>::s[t];
                ax += sx[r - 1] * c;
                ay += sy[r - 1] * c;
                bx += dx[r - 1] * sn;
                by += dy[r - 1] * sn;
            }
            // forward: X_q = A - i B, X_(R-q) = A + i B; inverse the other way round (i B = (-by, bx))
            const double sgn = inv ? -1.0 : 1.0;
            out[base + q * Ns].x = ax + sgn * by;
            out[base + q * Ns].y = ay - sgn * bx;
            out[base + (R - q) * Ns].x = ax - sgn * by;
            out[base + (R - q) * Ns].y = ay + sgn * bx;
        }
    }
}

// j / d for 0 <= j < 2^24 and small d, from a float reciprocal and one correction each way
__device__ __forceinline__ int div_small(int j, int d, float inv_d) {
    int q = (int)((float)j * inv_d);
    q -= q * d > j ? 1 : 0;
    q += (q + 1) * d <= j ? 1 : 0;
    return q;
}

// Transform the line of this thread's group (P = 2^lgP threads per line,
// line = threadIdx.x >> lgP) held in a; b is scratch.  Every thread of the
// block calls it (active = the line exists); returns the buffer holding the
// result, with the block synchronised.
__device__ double2* fft_line(double2* a, double2* b, bool active, const FftAxis& ax, bool inv) {
    const int P = 1 << ax.lgP, lane = threadIdx.x & (P - 1);
    for (int s = 0; s < ax.nfac; ++s) {
        if (active) {
            const int R = ax.fac[s], Ns = ax.ns[s], nb = ax.nb[s], step = ax.step[s];
            const float inv_ns = ax.inv_ns[s];
            for (int j = lane; j < nb; j += P) {
                const int jq = Ns == 1 ? j : div_small(j, Ns, inv_ns);
                switch (R) {
                    case 2: butterfly<2>(a, b, j, jq, Ns, nb, step, ax.tw, inv); break;
                    case 3: butterfly<3>(a, b, j, jq, Ns, nb, step, ax.tw, inv); break;
                    case 4: butterfly<4>(a, b, j, jq, Ns, nb, step, ax.tw, inv); break;
                    case 5: butterfly<5>(a, b, j, jq, Ns, nb, step, ax.tw, inv); break;
                    case 7: butterfly<7>(a, b, j, jq, Ns, nb, step, ax.tw, inv); break;
                    case 11: butterfly<11>(a, b, j, jq, Ns, nb, step, ax.tw, inv); break;
                    default: butterfly<13>(a, b, j, jq, Ns, nb, step, ax.tw, inv); break;
                }
            }
        }
        __syncthreads();
        double2* t = a;
        a = b;
        b = t;
    }
    return a;
}

// In-place stages: every thread holds its butterflies in registers across a
// barrier, so a batch of lines needs a single LDS buffer (the ping-pong of
// fft_line needs two).  Used when every stage of an axis has at most
// kInplaceBpt butterflies per thread.
constexpr int kInplaceBpt = 1;

// DFT of one radix-R butterfly held in vx, vy (inputs in r order, outputs in q
// order), twiddled for column k of the stage as in butterfly<R>.
template <int R>
__device__ __forceinline__ void bf_compute(double (&vx)[R], double (&vy)[R], int k, int Ns, int step,
                                           const double2* __restrict__ tw, bool inv) {
    if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
            const double wx = tw[k * r * step].x, wy0 = tw[k * r * step].y;
            const double wy = inv ? -wy0 : wy0;
            const double x = vx[r] * wx - vy[r] * wy, y = vx[r] * wy + vy[r] * wx;
            vx[r] = x;
            vy[r] = y;
        }
    }
    if constexpr (R % 2 == 0) {
        double ox[R], oy[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            double ax = vx[0], ay = vy[0];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const int t = (r * q) % R;
                const double c = Roots<R>::c[t], sn = inv ? Roots<R>::s[t] : -Roots<R>::s[t];
                ax += vx[r] * c - vy[r] * sn;
                ay += vx[r] * sn + vy[r] * c;
            }
            ox[q] = ax;
            oy[q] = ay;
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
            vx[q] = ox[q];
            vy[q] = oy[q];
        }
    } else {
        constexpr int H = (R - 1) / 2;
        double sx[H], sy[H], dx[H], dy[H];
        double x0 = vx[0], y0 = vy[0];
#pragma unroll
        for (int r = 1; r <= H; ++r) {
            sx[r - 1] = vx[r] + vx[R - r];
            sy[r - 1] = vy[r] + vy[R - r];
            dx[r - 1] = vx[r] - vx[R - r];
            dy[r - 1] = vy[r] - vy[R - r];
            x0 += sx[r - 1];
            y0 += sy[r - 1];
        }
        const double v0x = vx[0], v0y = vy[0];
        vx[0] = x0;
        vy[0] = y0;
        const double sgn = inv ? -1.0 : 1.0;
#pragma unroll
        for (int q = 1; q <= H; ++q) {
            double ax = v0x, ay = v0y, bx = 0.0, by = 0.0;
#pragma unroll
            for (int r = 1; r <= H; ++r) {
                const int t = (r * q) % R;
                const double c = Roots<R>::c[t], sn = Roots<R>::s[t];
                ax += sx[r - 1] * c;
                ay += sy[r - 1] * c;
                bx += dx[r - 1] * sn;
                by += dy[r - 1] * sn;
            }
            vx[q] = ax + sgn * by;
            vy[q] = ay - sgn * bx;
            vx[R - q] = ax - sgn * by;
            vy[R - q] = ay + sgn * bx;
        }
    }
}

// Stage s of the lines buf[l * ld + (0..n-1)], l < L, in place: thread t owns
// butterflies t, t + blockDim.x, ... (L * nb <= kInplaceBpt * blockDim.x), held in
// registers across the barrier.  Called by every thread of the block.
template <int R>
__device__ __forceinline__ void stage_inplace(double2* buf, int L, int ld, const FftAxis& ax, int s, bool inv) {
    const int Ns = ax.ns[s], nb = ax.nb[s], step = ax.step[s];
    const float inv_nb = 1.0f / (float)nb;
    double vx[kInplaceBpt][R], vy[kInplaceBpt][R];
    int base[kInplaceBpt];
    bool act[kInplaceBpt];
#pragma unroll
    for (int t = 0; t < kInplaceBpt; ++t) {
        const int b = threadIdx.x + t * blockDim.x;
        act[t] = b < L * nb;
        base[t] = 0;
        if (act[t]) {
            const int l = div_small(b, nb, inv_nb), j = b - l * nb;
            const int jq = Ns == 1 ? j : div_small(j, Ns, ax.inv_ns[s]);
            const double2* line = buf + l * ld;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                vx[t][r] = line[j + r * nb].x;
                vy[t][r] = line[j + r * nb].y;
            }
            bf_compute<R>(vx[t], vy[t], j - jq * Ns, Ns, step, ax.tw, inv);
            base[t] = l * ld + jq * Ns * R + (j - jq * Ns);
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kInplaceBpt; ++t)
        if (act[t]) {
#pragma unroll
            for (int q = 0; q < R; ++q) {
                buf[base[t] + q * Ns].x = vx[t][q];
                buf[base[t] + q * Ns].y = vy[t][q];
            }
        }
    __syncthreads();
}

__device__ void fft_lines_inplace(double2* buf, int L, int ld, const FftAxis& ax, bool inv) {
    for (int s = 0; s < ax.nfac; ++s) {
        switch (ax.fac[s]) {  // uniform over the block: the barriers inside are reached by all
            case 2: stage_inplace<2>(buf, L, ld, ax, s, inv); break;
            case 3: stage_inplace<3>(buf, L, ld, ax, s, inv); break;
            case 4: stage_inplace<4>(buf, L, ld, ax, s, inv); break;
            case 5: stage_inplace<5>(buf, L, ld, ax, s, inv); break;
            case 7: stage_inplace<7>(buf, L, ld, ax, s, inv); break;
            case 11: stage_inplace<11>(buf, L, ld, ax, s, inv); break;
            default: stage_inplace<13>(buf, L, ld, ax, s, inv); break;
        }
    }
}

// z pass, forward, after the block's lines are in LDS (lds[l * N + z], l < m):
// transform them and store their half spectra
template <int C>
__device__ __forceinline__ void z_r2c_lines(double2* lds, double2* __restrict__ H, int L0, int m, const FftAxis& az,
                                            int Kh, int Khp) {
    const int N = az.n, M = kFftThreads >> az.lgP, P = 1 << az.lgP;
    const int l = threadIdx.x >> az.lgP, lane = threadIdx.x & (P - 1);
    const double2* r = lds + l * N;
    if (az.inplace)
        fft_lines_inplace(lds, m, N, az, false);
    else
        r = fft_line(lds + l * N, lds + (M + l) * N, l < m, az, false);
    if (l < m) {
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int k = lane + u * P;
            if (k < Kh) {
                H[(size_t)(L0 + l) * Khp + k].x = r[k].x;
                H[(size_t)(L0 + l) * Khp + k].y = r[k].y;
            }
        }
    }
}

// z pass, forward: lines of the real grid -> half spectrum
template <int C>
__global__ void __launch_bounds__(kFftThreads) k_fft_z_r2c(const double* __restrict__ real, double2* __restrict__ H,
                                                           int nlines, FftAxis az, int Kh, int Khp,
                                                           const int* __restrict__ stop) {
    critical_path_wave();
    extern __shared__ double2 lds[];
    if (stop && *stop) return;
    const int N = az.n, M = kFftThreads >> az.lgP;
    const int L0 = blockIdx.x * M, m = min(M, nlines - L0);
    {  // all loads in flight before the LDS stores
        double v[C];
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int e = threadIdx.x + u * kFftThreads;
            if (e < m * N) v[u] = real[(size_t)L0 * N + e];
        }
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int e = threadIdx.x + u * kFftThreads;
            if (e < m * N) lds[e] = make_double2(v[u], 0.0);
        }
    }
    __syncthreads();
    z_r2c_lines<C>(lds, H, L0, m, az, Kh, Khp);
}

// Dipole spreading and the forward z pass in one kernel: the block's grid
// points (whole z lines) are summed from their rows of the spreading table
// (rowstart / esite / ecoef, entries in site order, the same sums as a
// separate spreading pass) straight into LDS, so the real grid is neither
// written nor read back.  A thread's points advance through their rows in
// lockstep, so that the loads of up to C entries are in flight together.
template <int C>
__global__ void __launch_bounds__(kFftThreads) k_spread_fft_z_r2c(const int* __restrict__ rowstart,
                                                                  const int* __restrict__ esite,
                                                                  const double* __restrict__ ecoef,
                                                                  const double* __restrict__ mu, double2* __restrict__ H,
                                                                  int nlines, FftAxis az, int Kh, int Khp,
                                                                  const int* __restrict__ stop) {
    critical_path_wave();
    extern __shared__ double2 lds[];
    if (stop && *stop) return;
    const int N = az.n, M = kFftThreads >> az.lgP;
    const int L0 = blockIdx.x * M, m = min(M, nlines - L0);
    {
        int rb[C], re[C];
        double v[C];
        int most = 0;
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int e = threadIdx.x + u * kFftThreads;
            const size_t g = (size_t)L0 * N + e;
            rb[u] = e < m * N ? rowstart[g] : 0;
            re[u] = e < m * N ? rowstart[g + 1] : 0;
            v[u] = 0.0;
            most = max(most, re[u] - rb[u]);
        }
        for (int k = 0; k < most; ++k) {
#pragma unroll
            for (int u = 0; u < C; ++u) {
                const int e = rb[u] + k;
                if (e < re[u]) {
                    const int s = esite[e];
                    v[u] += mu[3 * s] * ecoef[3 * (size_t)e] + mu[3 * s + 1] * ecoef[3 * (size_t)e + 1] +
                            mu[3 * s + 2] * ecoef[3 * (size_t)e + 2];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int e = threadIdx.x + u * kFftThreads;
            if (e < m * N) lds[e] = make_double2(v[u], 0.0);
        }
    }
    __syncthreads();
    z_r2c_lines<C>(lds, H, L0, m, az, Kh, Khp);
}

// y pass (either direction): block = (x, tile of M consecutive kz)
template <int C>
__global__ void __launch_bounds__(kFftThreads) k_fft_y(double2* __restrict__ H, int K1, int Kh, int Khp, FftAxis ay,
                                                       int inv, const int* __restrict__ stop) {
    critical_path_wave();
    extern __shared__ double2 lds[];
    if (stop && *stop) return;
    const int N = ay.n, lgM = 7 - ay.lgP, M = 1 << lgM;
    const int ntile = (Kh + M - 1) >> lgM;
    const int x = blockIdx.x / ntile, kz0 = (blockIdx.x - x * ntile) * M, m = min(M, Kh - kz0);
    const int cl = threadIdx.x & (M - 1), cy = threadIdx.x >> lgM;  // copy mapping: consecutive kz per row
    const int ystep = kFftThreads >> lgM;
    if (cl < m) {
        double vx[C], vy[C];
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int y = cy + u * ystep;
            if (y < N) {
                const double2* src = H + ((size_t)x * K1 + y) * Khp + kz0 + cl;
                vx[u] = src->x;
                vy[u] = src->y;
            }
        }
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int y = cy + u * ystep;
            if (y < N) {
                lds[cl * N + y].x = vx[u];
                lds[cl * N + y].y = vy[u];
            }
        }
    }
    __syncthreads();
    const int l = threadIdx.x >> ay.lgP;
    int off = 0;  // 0 or M N: which buffer holds the result
    if (ay.inplace)
        fft_lines_inplace(lds, m, N, ay, inv != 0);
    else
        off = (int)(fft_line(lds + l * N, lds + (M + l) * N, l < m, ay, inv != 0) - (lds + l * N));
    if (cl < m) {
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int y = cy + u * ystep;
            if (y < N) {
                double2* dst = H + ((size_t)x * K1 + y) * Khp + kz0 + cl;
                dst->x = lds[off + cl * N + y].x;
                dst->y = lds[off + cl * N + y].y;
            }
        }
    }
}

// x pass: forward, influence function, inverse; block = (y, tile of M consecutive kz)
template <int C>
__global__ void __launch_bounds__(kFftThreads) k_fft_x_conv(double2* __restrict__ H, const double* __restrict__ influence,
                                                            int K1, int Kh, int Khp, FftAxis ax,
                                                            const int* __restrict__ stop) {
    critical_path_wave();
    extern __shared__ double2 lds[];
    if (stop && *stop) return;
    const int N = ax.n, lgM = 7 - ax.lgP, M = 1 << lgM, P = 1 << ax.lgP;
    const int ntile = (Kh + M - 1) >> lgM;
    const int y = blockIdx.x / ntile, kz0 = (blockIdx.x - y * ntile) * M, m = min(M, Kh - kz0);
    const int cl = threadIdx.x & (M - 1), cx = threadIdx.x >> lgM;
    const int xstep = kFftThreads >> lgM;
    const int l = threadIdx.x >> ax.lgP, lane = threadIdx.x & (P - 1);
    double f[C];  // influence factors of this thread's k-space points, fetched early
#pragma unroll
    for (int u = 0; u < C; ++u) {
        const int kx = lane + u * P;
        f[u] = (l < m && kx < N) ? influence[((size_t)kx * K1 + y) * Khp + kz0 + l] : 0.0;
    }
    if (cl < m) {
        double vx[C], vy[C];
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int x = cx + u * xstep;
            if (x < N) {
                const double2* src = H + ((size_t)x * K1 + y) * Khp + kz0 + cl;
                vx[u] = src->x;
                vy[u] = src->y;
            }
        }
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int x = cx + u * xstep;
            if (x < N) {
                lds[cl * N + x].x = vx[u];
                lds[cl * N + x].y = vy[u];
            }
        }
    }
    __syncthreads();
    double2* a = lds + l * N;
    double2* b = lds + (M + l) * N;
    double2* r = a;
    if (ax.inplace)
        fft_lines_inplace(lds, m, N, ax, false);
    else
        r = fft_line(a, b, l < m, ax, false);
    if (l < m) {
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int kx = lane + u * P;
            if (kx < N) {
                r[kx].x = r[kx].x * f[u];
                r[kx].y = r[kx].y * f[u];
            }
        }
    }
    __syncthreads();
    if (ax.inplace)
        fft_lines_inplace(lds, m, N, ax, true);
    else
        r = fft_line(r, r == a ? b : a, l < m, ax, true);
    const int off = (int)(r - (lds + l * N));
    if (cl < m) {
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int x = cx + u * xstep;
            if (x < N) {
                double2* dst = H + ((size_t)x * K1 + y) * Khp + kz0 + cl;
                dst->x = lds[off + cl * N + x].x;
                dst->y = lds[off + cl * N + x].y;
            }
        }
    }
}

// z pass, inverse: half spectrum -> real lines (kz >= Kh from conj(H[N - kz]))
template <int C>
__global__ void __launch_bounds__(kFftThreads) k_fft_z_c2r(const double2* __restrict__ H, double* __restrict__ real,
                                                           int nlines, FftAxis az, int Kh, int Khp,
                                                           const int* __restrict__ stop) {
    critical_path_wave();
    extern __shared__ double2 lds[];
    if (stop && *stop) return;
    const int N = az.n, M = kFftThreads >> az.lgP, P = 1 << az.lgP;
    const int L0 = blockIdx.x * M, m = min(M, nlines - L0);
    const int l = threadIdx.x >> az.lgP, lane = threadIdx.x & (P - 1);
    double2* a = lds + l * N;
    if (l < m) {
        const size_t row = (size_t)(L0 + l) * Khp;
        double vx[C], vy[C];
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int k = lane + u * P;
            if (k < N) {
                const double2* src = H + row + (k < Kh ? k : N - k);
                vx[u] = src->x;
                vy[u] = src->y;
            }
        }
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int k = lane + u * P;
            if (k < N) {
                a[k].x = vx[u];
                a[k].y = k < Kh ? vy[u] : -vy[u];
            }
        }
    }
    __syncthreads();
    const double2* r = a;
    if (az.inplace)
        fft_lines_inplace(lds, m, N, az, true);
    else
        r = fft_line(a, lds + (M + l) * N, l < m, az, true);
    const int off = (int)(r - a);
#pragma unroll
    for (int u = 0; u < C; ++u) {
        const int e = threadIdx.x + u * kFftThreads;
        if (e < m * N) real[(size_t)L0 * N + e] = lds[off + e].x;
    }
}

// ---------------------------------------------------------------------------
// Plane passes: when one x plane of the grid fits in LDS, the z and y passes
// of each direction run as one kernel per plane (one workgroup per x), so the
// half spectrum makes one round trip through HBM less per direction.  Two real
// z lines (y = 2p, 2p + 1) share one complex transform: Z = r_2p + i r_2p+1,
// separated afterwards with X_2p[k] = (Z[k] + conj Z[-k]) / 2 and
// X_2p+1[k] = (Z[k] - conj Z[-k]) / 2i (and the reverse on the way back).
// Every stage works in place, so the plane needs a single LDS buffer.
// ---------------------------------------------------------------------------
// A plane workgroup takes a whole CU (16 waves, ~118 KB of LDS at 121 x 121),
// so the plane passes only pay when nothing else runs beside them.
constexpr int kPlaneThreads = 1024;
constexpr int kPlaneCopies = 8;  // plane entries per thread in the copy phases (<= 8192)
constexpr int kPlanePairs = 4;   // (k, -k) pairs per thread in the separation (<= 4096)

// Forward, plane x = blockIdx.x: real[x][y][z] -> z transform (two lines per
// complex transform) -> y transform -> H[x][ky][kz], kz < Kh.
// LDS: packed lines A[p][z] (p < NP = ceil(K1/2)), then B[kz][y].
__global__ void __launch_bounds__(kPlaneThreads) k_fft_plane_fwd(const double* __restrict__ real, double2* __restrict__ H,
                                                                 int K1, int Kh, int Khp, FftAxis ay, FftAxis az,
                                                                 const int* __restrict__ stop) {
    critical_path_wave();
    extern __shared__ double2 lds[];
    if (stop && *stop) return;
    const int K2 = az.n, NP = (K1 + 1) >> 1, x = blockIdx.x;
    const double* src = real + (size_t)x * K1 * K2;
    {  // packed lines, all loads in flight before the LDS stores
        double re[kPlaneCopies], im[kPlaneCopies];
#pragma unroll
        for (int u = 0; u < kPlaneCopies; ++u) {
            const int e = threadIdx.x + u * kPlaneThreads;
            const int p = e / K2, z = e - p * K2;
            re[u] = e < NP * K2 ? src[(size_t)(2 * p) * K2 + z] : 0.0;
            im[u] = e < NP * K2 && 2 * p + 1 < K1 ? src[(size_t)(2 * p + 1) * K2 + z] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kPlaneCopies; ++u) {
            const int e = threadIdx.x + u * kPlaneThreads;
            if (e < NP * K2) lds[e] = make_double2(re[u], im[u]);
        }
    }
    __syncthreads();
    fft_lines_inplace(lds, NP, K2, az, false);
    {  // separate the two lines of each transform into B[kz][y]
        double zx[kPlanePairs], zy[kPlanePairs], cx[kPlanePairs], cy[kPlanePairs];
#pragma unroll
        for (int u = 0; u < kPlanePairs; ++u) {
            const int e = threadIdx.x + u * kPlaneThreads;
            if (e < NP * Kh) {
                const int p = e / Kh, k = e - p * Kh;
                const double2 a = lds[p * K2 + k], b = lds[p * K2 + (k == 0 ? 0 : K2 - k)];
                zx[u] = a.x;
                zy[u] = a.y;
                cx[u] = b.x;
                cy[u] = -b.y;  // conj Z[-k]
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kPlanePairs; ++u) {
            const int e = threadIdx.x + u * kPlaneThreads;
            if (e < NP * Kh) {
                const int p = e / Kh, k = e - p * Kh;
                // X_2p = (Z + conj Z-) / 2; X_2p+1 = (Z - conj Z-) / 2i = (Im, -Re) of (Z - conj Z-) / 2
                lds[k * K1 + 2 * p] = make_double2(0.5 * (zx[u] + cx[u]), 0.5 * (zy[u] + cy[u]));
                if (2 * p + 1 < K1) lds[k * K1 + 2 * p + 1] = make_double2(0.5 * (zy[u] - cy[u]), -0.5 * (zx[u] - cx[u]));
            }
        }
    }
    __syncthreads();
    fft_lines_inplace(lds, Kh, K1, ay, false);
    double2* dst = H + (size_t)x * K1 * Khp;
#pragma unroll
    for (int u = 0; u < kPlaneCopies; ++u) {
        const int e = threadIdx.x + u * kPlaneThreads;
        if (e < K1 * Kh) {
            const int y = e / Kh, k = e - y * Kh;
            const double2 v = lds[k * K1 + y];
            dst[(size_t)y * Khp + k].x = v.x;
            dst[(size_t)y * Khp + k].y = v.y;
        }
    }
}

// Inverse, plane x: H[x][ky][kz] -> y transform -> z transform of the
// Hermitian-extended lines (two per complex transform) -> real[x][y][z].
__global__ void __launch_bounds__(kPlaneThreads) k_fft_plane_inv(const double2* __restrict__ H, double* __restrict__ real,
                                                                 int K1, int Kh, int Khp, FftAxis ay, FftAxis az,
                                                                 const int* __restrict__ stop) {
    critical_path_wave();
    extern __shared__ double2 lds[];
    if (stop && *stop) return;
    const int K2 = az.n, NP = (K1 + 1) >> 1, x = blockIdx.x;
    const double2* src = H + (size_t)x * K1 * Khp;
    {
        double vx[kPlaneCopies], vy[kPlaneCopies];
#pragma unroll
        for (int u = 0; u < kPlaneCopies; ++u) {
            const int e = threadIdx.x + u * kPlaneThreads;
            const int y = e / Kh, k = e - y * Kh;
            if (e < K1 * Kh) {
                vx[u] = src[(size_t)y * Khp + k].x;
                vy[u] = src[(size_t)y * Khp + k].y;
            }
        }
#pragma unroll
        for (int u = 0; u < kPlaneCopies; ++u) {
            const int e = threadIdx.x + u * kPlaneThreads;
            const int y = e / Kh, k = e - y * Kh;
            if (e < K1 * Kh) lds[k * K1 + y] = make_double2(vx[u], vy[u]);
        }
    }
    __syncthreads();
    fft_lines_inplace(lds, Kh, K1, ay, true);
    {  // Z[p][k] = S_2p[k] + i S_2p+1[k], S_y[k] = B[k][y] (k < Kh) or conj B[K2-k][y]
        double zx[kPlaneCopies], zy[kPlaneCopies];
#pragma unroll
        for (int u = 0; u < kPlaneCopies; ++u) {
            const int e = threadIdx.x + u * kPlaneThreads;
            if (e < NP * K2) {
                const int p = e / K2, k = e - p * K2;
                const bool lo = k < Kh;
                const int kk = lo ? k : K2 - k;
                const double2 a = lds[kk * K1 + 2 * p];
                const double2 b = 2 * p + 1 < K1 ? lds[kk * K1 + 2 * p + 1] : make_double2(0.0, 0.0);
                const double ay_ = lo ? a.y : -a.y, bx = b.x, by = lo ? b.y : -b.y;
                zx[u] = a.x - by;
                zy[u] = ay_ + bx;
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kPlaneCopies; ++u) {
            const int e = threadIdx.x + u * kPlaneThreads;
            if (e < NP * K2) lds[e] = make_double2(zx[u], zy[u]);
        }
    }
    __syncthreads();
    fft_lines_inplace(lds, NP, K2, az, true);
    double* dst = real + (size_t)x * K1 * K2;
#pragma unroll
    for (int u = 0; u < kPlaneCopies; ++u) {
        const int e = threadIdx.x + u * kPlaneThreads;
        if (e < NP * K2) {
            const int p = e / K2, z = e - p * K2;
            const double2 v = lds[e];
            dst[(size_t)(2 * p) * K2 + z] = v.x;
            if (2 * p + 1 < K1) dst[(size_t)(2 * p + 1) * K2 + z] = v.y;
        }
    }
}

}  // namespace

static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

// LDS entries of one plane: packed z lines and y lines of the half spectrum
static size_t plane_entries(const PmeFft& F) {
    const size_t K1 = F.ax[1].n, K2 = F.ax[2].n, Kh = K2 / 2 + 1, NP = (K1 + 1) / 2;
    return std::max(NP * K2, Kh * K1);
}

bool fft_plane_ok(const PmeFft& F) {
    const int K1 = F.ax[1].n, K2 = F.ax[2].n, Kh = K2 / 2 + 1, NP = (K1 + 1) / 2;
    if (plane_entries(F) > (size_t)kPlaneCopies * kPlaneThreads) return false;
    if ((size_t)NP * Kh > (size_t)kPlanePairs * kPlaneThreads) return false;
    if (plane_entries(F) * sizeof(double2) > 160 * 1024) return false;
    // every stage of both axes within kInplaceBpt butterflies per thread
    for (int s = 0; s < F.ax[2].nfac; ++s)
        if (NP * F.ax[2].nb[s] > kInplaceBpt * kPlaneThreads) return false;
    for (int s = 0; s < F.ax[1].nfac; ++s)
        if (Kh * F.ax[1].nb[s] > kInplaceBpt * kPlaneThreads) return false;
    return true;
}

void launch_fft_plane_fwd(hipStream_t s, const PmeFft& F, const double* real, double2* H, const KEvents& ev,
                          const int* stop) {
    const int K1 = F.ax[1].n, Kh = F.ax[2].n / 2 + 1;
    static const hipError_t attr = hipFuncSetAttribute((const void*)k_fft_plane_fwd,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)attr;  // a failed launch shows as a launch error
    MBPOL_LAUNCH(ev, k_fft_plane_fwd, F.ax[0].n, kPlaneThreads, plane_entries(F) * sizeof(double2), s, real, H, K1, Kh,
                 F.khp, F.ax[1], F.ax[2], stop);
}
void launch_fft_plane_inv(hipStream_t s, const PmeFft& F, const double2* H, double* real, const KEvents& ev,
                          const int* stop) {
    const int K1 = F.ax[1].n, Kh = F.ax[2].n / 2 + 1;
    static const hipError_t attr = hipFuncSetAttribute((const void*)k_fft_plane_inv,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)attr;
    MBPOL_LAUNCH(ev, k_fft_plane_inv, F.ax[0].n, kPlaneThreads, plane_entries(F) * sizeof(double2), s, H, real, K1, Kh,
                 F.khp, F.ax[1], F.ax[2], stop);
}

// C: elements a thread moves per copy phase, >= ceil(n / P) (at most 13 for the supported lengths)
static int copies(const FftAxis& a) {
    const int c = (a.n + (1 << a.lgP) - 1) >> a.lgP;
    return c <= 4 ? 4 : c <= 8 ? 8 : 16;
}
#define MBPOL_FFT_LAUNCH(kern, axis, ...)                        \
    switch (copies(axis)) {                                      \
        case 4: MBPOL_LAUNCH(ev, kern<4>, __VA_ARGS__); break;   \
        case 8: MBPOL_LAUNCH(ev, kern<8>, __VA_ARGS__); break;   \
        default: MBPOL_LAUNCH(ev, kern<16>, __VA_ARGS__); break; \
    }

// one buffer of M lines for the in-place stages, two for the ping-pong ones
static size_t lds_bytes(const FftAxis& a) {
    return (a.inplace ? 1 : 2) * sizeof(double2) * (size_t)(kFftThreads >> a.lgP) * a.n;
}

void launch_fft_z_r2c(hipStream_t s, const PmeFft& F, const double* real, double2* H, const KEvents& ev,
                      const int* stop) {
    const int nl = F.ax[0].n * F.ax[1].n, M = kFftThreads >> F.ax[2].lgP, Kh = F.ax[2].n / 2 + 1;
    MBPOL_FFT_LAUNCH(k_fft_z_r2c, F.ax[2], nblk(nl, M), kFftThreads, lds_bytes(F.ax[2]), s, real, H, nl, F.ax[2], Kh,
                     F.khp, stop)
}
void launch_spread_fft_z_r2c(hipStream_t s, const PmeFft& F, const int* rowstart, const int* esite,
                             const double* ecoef, const double* mu, double2* H, const KEvents& ev, const int* stop) {
    const int nl = F.ax[0].n * F.ax[1].n, M = kFftThreads >> F.ax[2].lgP, Kh = F.ax[2].n / 2 + 1;
    MBPOL_FFT_LAUNCH(k_spread_fft_z_r2c, F.ax[2], nblk(nl, M), kFftThreads, lds_bytes(F.ax[2]), s, rowstart, esite,
                     ecoef, mu, H, nl, F.ax[2], Kh, F.khp, stop)
}
void launch_fft_y(hipStream_t s, const PmeFft& F, double2* H, int inverse, const KEvents& ev, const int* stop) {
    const int K0 = F.ax[0].n, K1 = F.ax[1].n, Kh = F.ax[2].n / 2 + 1, M = kFftThreads >> F.ax[1].lgP;
    MBPOL_FFT_LAUNCH(k_fft_y, F.ax[1], K0 * nblk(Kh, M), kFftThreads, lds_bytes(F.ax[1]), s, H, K1, Kh, F.khp, F.ax[1],
                     inverse, stop)
}
void launch_fft_x_conv(hipStream_t s, const PmeFft& F, double2* H, const double* influence, const KEvents& ev,
                       const int* stop) {
    const int K1 = F.ax[1].n, Kh = F.ax[2].n / 2 + 1, M = kFftThreads >> F.ax[0].lgP;
    MBPOL_FFT_LAUNCH(k_fft_x_conv, F.ax[0], K1 * nblk(Kh, M), kFftThreads, lds_bytes(F.ax[0]), s, H, influence, K1, Kh,
                     F.khp, F.ax[0], stop)
}
void launch_fft_z_c2r(hipStream_t s, const PmeFft& F, const double2* H, double* real, const KEvents& ev,
                      const int* stop) {
    const int nl = F.ax[0].n * F.ax[1].n, M = kFftThreads >> F.ax[2].lgP, Kh = F.ax[2].n / 2 + 1;
    MBPOL_FFT_LAUNCH(k_fft_z_c2r, F.ax[2], nblk(nl, M), kFftThreads, lds_bytes(F.ax[2]), s, H, real, nl, F.ax[2], Kh,
                     F.khp, stop)
}
#undef MBPOL_FFT_LAUNCH

}  // namespace mbpol
