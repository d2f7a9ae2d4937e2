// Learned initial guess of the induced dipoles, and Langevin dynamics to
// produce its training data.  No counterpart in the reference, which always
// starts the dipole iteration from alpha * E_fixed
// (MBPolReferenceElectrostaticsForce.cpp:628-639).
//
// Model: for every polarisable atom (O, H1, H2 of each water; sites 4w..4w+2 of
// the electrostatics layout), 48 radial features -- 16 Gaussians on
// 0.08..0.6 nm times 3 neighbour channels (O, H, M sites), with a cosine cutoff
// at 0.6 nm -- pass through a 48 -> 64 -> 64 -> 3 network with SiLU
// activations.  The 3 outputs are a correction to a base guess relative to the
// length of an anchor vector, in a frame of the atom: e1 along the anchor, e2
// along the part of a molecular direction orthogonal to e1 (the HOH bisector
// for O, O -> H for an H), e3 = e1 x e2; in units of a scale fixed at training
// time.  From alpha E_fixed: base = anchor = alpha E_fixed, i.e. the outputs are
// mu - alpha E_fixed relative to |alpha E_fixed| (at least kGuessFloor).  As a
// correction to the previous step's dipoles: base = those dipoles, anchor =
// the change of alpha E_fixed since that step.  Mutual induction mostly amplifies the fixed-
// field dipole by a factor that depends on the surroundings, which is what the
// radial features carry.  The matrix products run on the BF16 MFMA units
// (v_mfma_f32_32x32x16_bf16) with FP32 accumulation; biases, activations and
// the last layer are FP32.
#include "device_common.h"
#include "guess_kernels.h"

namespace mbpol {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

constexpr float kRbfLo = 0.08f, kRbfHi = 0.6f;
constexpr float kRbfStep = (kRbfHi - kRbfLo) / (kGuessRbf - 1);
constexpr float kRbfEta = 1.0f / (2.0f * kRbfStep * kRbfStep);

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ void mimage(double& dx, double& dy, double& dz, const GuessGeom& G) {
    if (!G.periodic) return;
    dx -= floor(dx * G.inv_box.x + 0.5) * G.box.x;
    dy -= floor(dy * G.inv_box.y + 0.5) * G.box.y;
    dz -= floor(dz * G.inv_box.z + 0.5) * G.box.z;
}

// ---------------------------------------------------------------------------
// features and frames: kGuessLanes lanes per atom, each accumulating the 48
// sums over its share of the candidates, then a group reduction
// ---------------------------------------------------------------------------
constexpr int kGuessLanes = 16;

__global__ void __launch_bounds__(256) k_guess_features(int nw, const double* __restrict__ pos, const int* __restrict__ type,
                                                        const double* __restrict__ anchor, GuessGeom G,
                                                        const int* __restrict__ cell_of,
                                                        const int* __restrict__ cell_start,
                                                        const int* __restrict__ members, float* __restrict__ feat,
                                                        float* __restrict__ frame) {
    const int a = (blockIdx.x * blockDim.x + threadIdx.x) / kGuessLanes, lane = threadIdx.x % kGuessLanes;
    if (a >= 3 * nw) return;  // uniform over the group
    const int w = a / 3, k = a % 3, s = 4 * w + k, nsite = 4 * nw;
    const double xs = pos[3 * s], ys = pos[3 * s + 1], zs = pos[3 * s + 2];
    float acc[3][kGuessRbf];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < kGuessRbf; ++r) acc[c][r] = 0.f;
    const double cut2 = (double)kRbfHi * kRbfHi;
    auto consider = [&](int t) {
        if (t == s) return;
        double dx = pos[3 * t] - xs, dy = pos[3 * t + 1] - ys, dz = pos[3 * t + 2] - zs;
        mimage(dx, dy, dz, G);
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 >= cut2) return;
        const float r = (float)sqrt(r2);
        const float fc = 0.5f * (1.0f + __cosf(3.14159265358979f * r / kRbfHi));
        const int c = type[t];
#pragma unroll
        for (int q = 0; q < kGuessRbf; ++q) {
            const float u = r - (kRbfLo + q * kRbfStep);
            const float g = fc * __expf(-kRbfEta * u * u);
            acc[0][q] += c == 0 ? g : 0.f;
            acc[1][q] += c == 1 ? g : 0.f;
            acc[2][q] += c == 2 ? g : 0.f;
        }
    };
    if (!G.use_cells) {
        for (int t = lane; t < nsite; t += kGuessLanes) consider(t);
    } else {
        const int cell = cell_of[s];
        const int cz = cell % G.nc.z, cy = (cell / G.nc.z) % G.nc.y, cx = cell / (G.nc.z * G.nc.y);
        for (int ax = -G.reach; ax <= G.reach; ++ax)
            for (int ay = -G.reach; ay <= G.reach; ++ay)
                for (int az = -G.reach; az <= G.reach; ++az) {
                    int X = cx + ax, Y = cy + ay, Z = cz + az;
                    X = X < 0 ? X + G.nc.x : (X >= G.nc.x ? X - G.nc.x : X);
                    Y = Y < 0 ? Y + G.nc.y : (Y >= G.nc.y ? Y - G.nc.y : Y);
                    Z = Z < 0 ? Z + G.nc.z : (Z >= G.nc.z ? Z - G.nc.z : Z);
                    const int cc = (X * G.nc.y + Y) * G.nc.z + Z;
                    for (int e = cell_start[cc] + lane; e < cell_start[cc + 1]; e += kGuessLanes) consider(members[e]);
                }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int q = 0; q < kGuessRbf; ++q) {
            float v = acc[c][q];
#pragma unroll
            for (int off = kGuessLanes / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kGuessLanes);
            acc[c][q] = v;
        }
    if (lane == 0) {
        float* f = feat + (size_t)a * kGuessFeat;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int q = 0; q < kGuessRbf; ++q) f[c * kGuessRbf + q] = acc[c][q];
    } else if (lane == 1) {
        // the atom's frame: e1 along the anchor vector, e2 from a molecular direction
        const int o = 4 * w;
        double ref[3];
        if (k == 0) {
            for (int d = 0; d < 3; ++d) ref[d] = 0.0;
            for (int m = 1; m <= 2; ++m) {
                double dv[3] = {pos[3 * (o + m)] - pos[3 * o], pos[3 * (o + m) + 1] - pos[3 * o + 1],
                                pos[3 * (o + m) + 2] - pos[3 * o + 2]};
                mimage(dv[0], dv[1], dv[2], G);
                for (int d = 0; d < 3; ++d) ref[d] += dv[d];
            }
        } else {
            ref[0] = pos[3 * s] - pos[3 * o];
            ref[1] = pos[3 * s + 1] - pos[3 * o + 1];
            ref[2] = pos[3 * s + 2] - pos[3 * o + 2];
            mimage(ref[0], ref[1], ref[2], G);
        }
        double e1[3] = {anchor[3 * s], anchor[3 * s + 1], anchor[3 * s + 2]}, e2[3], e3[3];
        double nrm = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        if (!(nrm > 1e-12)) {  // no field: the molecular direction takes its place
            for (int d = 0; d < 3; ++d) e1[d] = ref[d];
            nrm = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
            ref[0] += 1.0;  // any direction not along it
        }
        for (int d = 0; d < 3; ++d) e1[d] /= nrm;
        double re = ref[0] * e1[0] + ref[1] * e1[1] + ref[2] * e1[2];
        for (int d = 0; d < 3; ++d) e2[d] = ref[d] - re * e1[d];
        nrm = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
        if (!(nrm > 1e-9 * sqrt(ref[0] * ref[0] + ref[1] * ref[1] + ref[2] * ref[2]))) {  // parallel: any normal
            const double t[3] = {fabs(e1[0]) < 0.9 ? 1.0 : 0.0, fabs(e1[0]) < 0.9 ? 0.0 : 1.0, 0.0};
            re = t[0] * e1[0] + t[1] * e1[1];
            for (int d = 0; d < 3; ++d) e2[d] = t[d] - re * e1[d];
            nrm = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
        }
        for (int d = 0; d < 3; ++d) e2[d] /= nrm;
        e3[0] = e1[1] * e2[2] - e1[2] * e2[1];
        e3[1] = e1[2] * e2[0] - e1[0] * e2[2];
        e3[2] = e1[0] * e2[1] - e1[1] * e2[0];
        float* R = frame + (size_t)a * 9;
        for (int d = 0; d < 3; ++d) {
            R[d] = (float)e1[d];
            R[3 + d] = (float)e2[d];
            R[6 + d] = (float)e3[d];
        }
    }
}

// Length of the anchor vector of site s, the unit of the network's outputs:
// at least kGuessFloor, and, when a reference is given (the learned correction
// to the previous step's dipoles, anchored on the change of alpha E_fixed), at
// least 1% of the reference's length (alpha E_fixed now).
__device__ __forceinline__ double anchor_norm(const double* a, const double* ref, int s) {
    const double n = sqrt(a[3 * s] * a[3 * s] + a[3 * s + 1] * a[3 * s + 1] + a[3 * s + 2] * a[3 * s + 2]);
    double lo = kGuessFloor;
    if (ref) {
        const double r = sqrt(ref[3 * s] * ref[3 * s] + ref[3 * s + 1] * ref[3 * s + 1] + ref[3 * s + 2] * ref[3 * s + 2]);
        lo = 0.01 * r;
    }
    return fmax(n, lo);
}

// training targets: R (mu - base) / |anchor| per atom (base = anchor = alpha E_fixed for the
// guess from alpha E_fixed; base = the previous step's dipoles, anchor = the change of alpha
// E_fixed since then, for the correction to the previous step)
__global__ void k_guess_targets(int nw, const float* __restrict__ frame, const double* __restrict__ mu,
                                const double* __restrict__ base, const double* __restrict__ anchor,
                                const double* __restrict__ ref, float* __restrict__ targ) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= 3 * nw) return;
    const int s = 4 * (a / 3) + a % 3;
    const double inv = 1.0 / anchor_norm(anchor, ref, s);
    const double d[3] = {(mu[3 * s] - base[3 * s]) * inv, (mu[3 * s + 1] - base[3 * s + 1]) * inv,
                         (mu[3 * s + 2] - base[3 * s + 2]) * inv};
    const float* R = frame + (size_t)a * 9;
    for (int r = 0; r < 3; ++r) targ[(size_t)a * 3 + r] = (float)(R[3 * r] * d[0] + R[3 * r + 1] * d[1] + R[3 * r + 2] * d[2]);
}

// mu0 = base + |anchor| scale R^T y on the polarisable atoms of every water; other sites 0
__global__ void k_guess_apply(int nw, const float* __restrict__ frame, const float* __restrict__ y, float scale,
                              const double* __restrict__ base, const double* __restrict__ anchor,
                              const double* __restrict__ ref, double* __restrict__ mu0) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= 4 * nw) return;
    const int w = s / 4, k = s % 4;
    if (k == 3) {
        mu0[3 * s] = mu0[3 * s + 1] = mu0[3 * s + 2] = 0.0;
        return;
    }
    const int a = 3 * w + k;
    const float* R = frame + (size_t)a * 9;
    const float* ya = y + (size_t)a * 3;
    const double f = (double)scale * anchor_norm(anchor, ref, s);
    for (int d = 0; d < 3; ++d)
        mu0[3 * s + d] = base[3 * s + d] + f * ((double)R[d] * ya[0] + (double)R[3 + d] * ya[1] +
                                                (double)R[6 + d] * ya[2]);
}

// out = a - b, n3 values
__global__ void k_vec_sub(int n3, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) out[i] = a[i] - b[i];
}

// ---------------------------------------------------------------------------
// the network on the MFMA units: one wave per 32 samples, four waves per block
// MFMA 32x32x16 bf16 lane layout (lane l, r = l % 32, h = l / 32):
//   A operand: A[r][8h + q], B operand: B[8h + q][r], q = 0..7
//   accumulator register v: C[8 (v / 4) + 4 h + v % 4][r]
// ---------------------------------------------------------------------------
constexpr int kWaves = 4;

struct WaveLds {
    __bf16 x[32][kGuessFeat];  // normalised inputs
    __bf16 h1[32][kGuessHid];
    float h2[32][kGuessHid];
    float dy[32][4];
    __bf16 dz2[32][kGuessHid];
    __bf16 dz1[32][kGuessHid];
};

__device__ __forceinline__ int crow(int v, int h) { return 8 * (v / 4) + 4 * h + (v % 4); }

__device__ __forceinline__ f32x16 mfma(const bf16x8& a, const bf16x8& b, const f32x16& c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// mode 0: inference on rows [0, nrows): y = network(x) (unscaled outputs)
// mode 1: training on the samples idx[0 .. nrows): per-wave gradients of
//         (1/nrows) sum |y - t|^2 into grad[wave][kGuessParams], loss into lossp[wave]
template <int MODE>
__global__ void __launch_bounds__(64 * kWaves) k_mlp(int nrows, const int* __restrict__ idx, const float* __restrict__ feat,
                                                     const float* __restrict__ targ, GuessWeights W,
                                                     float* __restrict__ y, float* __restrict__ grad,
                                                     float* __restrict__ lossp) {
    __shared__ WaveLds lds_all[kWaves];
    const int wv = threadIdx.x / 64, l = threadIdx.x % 64, r = l % 32, h = l / 32;
    WaveLds& L = lds_all[wv];
    const int wave = blockIdx.x * kWaves + wv, row0 = wave * 32;
    // normalised inputs -> LDS (bf16); rows past the end are zero
    for (int e = l; e < 32 * kGuessFeat; e += 64) {
        const int i = e / kGuessFeat, f = e % kGuessFeat, rr = row0 + i;
        float v = 0.f;
        if (rr < nrows) {
            const size_t src = MODE == 1 ? (size_t)idx[rr] : (size_t)rr;
            v = (feat[src * kGuessFeat + f] - W.mean[f]) * W.inv_std[f];
        }
        L.x[i][f] = (__bf16)v;
    }
    __syncthreads();
    // layer 1: Z1 = X W1 + b1
    f32x16 z1[2];
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
        f32x16 c = {};
#pragma unroll
        for (int ks = 0; ks < kGuessFeat / 16; ++ks) {
            bf16x8 a, b;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                a[q] = L.x[r][16 * ks + 8 * h + q];
                b[q] = W.w1t[(32 * nb + r) * kGuessFeat + 16 * ks + 8 * h + q];
            }
            c = mfma(a, b, c);
        }
        const float bias = W.b1[32 * nb + r];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            c[v] += bias;
            L.h1[crow(v, h)][32 * nb + r] = (__bf16)silu(c[v]);
        }
        z1[nb] = c;
    }
    __syncthreads();
    // layer 2: Z2 = H1 W2 + b2
    f32x16 z2[2];
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
        f32x16 c = {};
#pragma unroll
        for (int ks = 0; ks < kGuessHid / 16; ++ks) {
            bf16x8 a, b;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                a[q] = L.h1[r][16 * ks + 8 * h + q];
                b[q] = W.w2t[(32 * nb + r) * kGuessHid + 16 * ks + 8 * h + q];
            }
            c = mfma(a, b, c);
        }
        const float bias = W.b2[32 * nb + r];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            c[v] += bias;
            L.h2[crow(v, h)][32 * nb + r] = silu(c[v]);
        }
        z2[nb] = c;
    }
    __syncthreads();
    // layer 3 (FP32): lane (r, h) sums half h of the hidden units for sample r
    float yo[3];
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        float s = 0.f;
        for (int j = 32 * h; j < 32 * h + 32; ++j) s += L.h2[r][j] * W.w3[j * 3 + o];
        s += __shfl_xor(s, 32, 64);
        yo[o] = s + W.b3[o];
    }
    const int rr = row0 + r;
    if (MODE == 0) {
        if (h == 0 && rr < nrows)
            for (int o = 0; o < 3; ++o) y[(size_t)rr * 3 + o] = yo[o];
        return;
    } else {
        // loss and dY = 2 (y - t) / nrows
        float ls = 0.f;
        if (h == 0) {
            for (int o = 0; o < 3; ++o) {
                float d = 0.f;
                if (rr < nrows) d = yo[o] - targ[(size_t)idx[rr] * 3 + o] * W.inv_tscale;
                ls += d * d;
                L.dy[r][o] = 2.f * d / (float)nrows;
            }
        }
        for (int off = 32; off > 0; off >>= 1) ls += __shfl_xor(ls, off, 64);
        float* G = grad + (size_t)wave * kGuessParams;
        if (l == 0) lossp[wave] = ls / (float)nrows;
        __syncthreads();
        // dW3 = H2^T dY, db3
        {
            const int j = l;
            for (int o = 0; o < 3; ++o) {
                float s = 0.f;
                for (int i = 0; i < 32; ++i) s += L.h2[i][j] * L.dy[i][o];
                G[kOffW3 + j * 3 + o] = s;
            }
            if (l < 3) {
                float s = 0.f;
                for (int i = 0; i < 32; ++i) s += L.dy[i][l];
                G[kOffB3 + l] = s;
            }
        }
        // dZ2 = (dY W3^T) * silu'(Z2), db2
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            const int j = 32 * nb + r;
            const float w30 = W.w3[j * 3], w31 = W.w3[j * 3 + 1], w32 = W.w3[j * 3 + 2];
            float sb = 0.f;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int i = crow(v, h);
                const float dz = (L.dy[i][0] * w30 + L.dy[i][1] * w31 + L.dy[i][2] * w32) * dsilu(z2[nb][v]);
                L.dz2[i][j] = (__bf16)dz;
                sb += dz;
            }
            sb += __shfl_xor(sb, 32, 64);
            if (h == 0) G[kOffB2 + j] = sb;
        }
        __syncthreads();
        // dW2[k][j] = sum_i H1[i][k] dZ2[i][j]: M = k, N = j, K = samples
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                f32x16 c = {};
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    bf16x8 a, b;
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        a[q] = L.h1[16 * s + 8 * h + q][32 * mb + r];
                        b[q] = L.dz2[16 * s + 8 * h + q][32 * nb + r];
                    }
                    c = mfma(a, b, c);
                }
#pragma unroll
                for (int v = 0; v < 16; ++v) G[kOffW2 + (32 * mb + crow(v, h)) * kGuessHid + 32 * nb + r] = c[v];
            }
        // dZ1 = (dZ2 W2^T) * silu'(Z1): M = samples, N = k, K = j; db1
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            f32x16 c = {};
#pragma unroll
            for (int ks = 0; ks < kGuessHid / 16; ++ks) {
                bf16x8 a, b;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    a[q] = L.dz2[r][16 * ks + 8 * h + q];
                    b[q] = W.w2[(32 * nb + r) * kGuessHid + 16 * ks + 8 * h + q];
                }
                c = mfma(a, b, c);
            }
            float sb = 0.f;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const float dz = c[v] * dsilu(z1[nb][v]);
                L.dz1[crow(v, h)][32 * nb + r] = (__bf16)dz;
                sb += dz;
            }
            sb += __shfl_xor(sb, 32, 64);
            if (h == 0) G[kOffB1 + 32 * nb + r] = sb;
        }
        __syncthreads();
        // dW1[f][k] = sum_i X[i][f] dZ1[i][k]: M = f (48 of 64), N = k, K = samples
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                f32x16 c = {};
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    bf16x8 a, b;
                    const int f = 32 * mb + r;
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        a[q] = f < kGuessFeat ? L.x[16 * s + 8 * h + q][f] : (__bf16)0.f;
                        b[q] = L.dz1[16 * s + 8 * h + q][32 * nb + r];
                    }
                    c = mfma(a, b, c);
                }
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int f = 32 * mb + crow(v, h);
                    if (f < kGuessFeat) G[kOffW1 + f * kGuessHid + 32 * nb + r] = c[v];
                }
            }
    }
}

// Adam step on the FP32 master parameters from the per-wave gradients (summed
// in wave order), then the BF16 operand copies; lossp summed into loss_out
__global__ void k_adam(int nwaves, const float* __restrict__ grad, const float* __restrict__ lossp, float* __restrict__ w,
                       float* __restrict__ m, float* __restrict__ v, int t, float lr, GuessWeights W,
                       float* __restrict__ loss_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        float s = 0.f;
        for (int k = 0; k < nwaves; ++k) s += lossp[k];
        *loss_out = s;
    }
    if (p >= kGuessParams) return;
    float g = 0.f;
    for (int k = 0; k < nwaves; ++k) g += grad[(size_t)k * kGuessParams + p];
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mm = b1 * m[p] + (1.f - b1) * g, vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.f - powf(b1, (float)t)), vh = vv / (1.f - powf(b2, (float)t));
    const float nw = w[p] - lr * mh / (sqrtf(vh) + eps);
    w[p] = nw;
    guess_publish(p, nw, W);
}

// BF16 operand copies of every parameter
__global__ void k_publish(const float* __restrict__ w, GuessWeights W) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kGuessParams) guess_publish(p, w[p], W);
}

// column sums (and sums of squares) of a row-major [nrows x ncol] float array
// over a contiguous chunk of rows per block, in double; ncol <= 256
__global__ void k_colsums(size_t nrows, int ncol, const float* __restrict__ a, size_t rows_per_block,
                          double* __restrict__ out) {
    const int c = threadIdx.x;
    if (c >= ncol) return;
    const size_t r0 = (size_t)blockIdx.x * rows_per_block, r1 = min(nrows, r0 + rows_per_block);
    double s = 0.0, s2 = 0.0;
    for (size_t rr = r0; rr < r1; ++rr) {
        const double x = a[rr * ncol + c];
        s += x;
        s2 += x * x;
    }
    out[((size_t)blockIdx.x * ncol + c) * 2] = s;
    out[((size_t)blockIdx.x * ncol + c) * 2 + 1] = s2;
}

// ---------------------------------------------------------------------------
// Langevin dynamics, BAOAB splitting (Leimkuhler and Matthews 2013): kicks with
// the forces (kJ/mol/nm) over masses (amu) give nm/ps^2; particles of zero
// inverse mass (virtual sites) never move on their own
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long splitmix(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// standard normal from a counter (Box-Muller on two 53-bit uniforms)
__device__ __forceinline__ double normal_at(unsigned long long seed, unsigned long long step, unsigned long long i) {
    const unsigned long long a = splitmix(seed ^ splitmix(step * 0x100000001B3ull + i));
    const unsigned long long b = splitmix(a);
    const double u1 = ((a >> 11) + 1.0) * (1.0 / 9007199254740993.0), u2 = (b >> 11) * (1.0 / 9007199254740992.0);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__global__ void k_md_kick(int n3, const double* __restrict__ f, const double* __restrict__ inv_mass, double h,
                          double* __restrict__ v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n3) v[t] += h * f[t] * inv_mass[t / 3];
}

__global__ void k_md_drift(int n3, const double* __restrict__ v, const double* __restrict__ inv_mass, double h,
                           double* __restrict__ x) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n3 && inv_mass[t / 3] != 0.0) x[t] += h * v[t];
}

// v = a v + sqrt((1 - a^2) kT / m) xi
__global__ void k_md_thermostat(int n3, const double* __restrict__ inv_mass, double a, double kT,
                                unsigned long long seed, unsigned long long step, double* __restrict__ v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n3) return;
    const double im = inv_mass[t / 3];
    v[t] = im == 0.0 ? 0.0 : a * v[t] + sqrt((1.0 - a * a) * kT * im) * normal_at(seed, step, (unsigned long long)t);
}

// kinetic energy, one block, fixed order
__global__ void k_md_kinetic(int n3, const double* __restrict__ v, const double* __restrict__ inv_mass,
                             double* __restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int t = threadIdx.x; t < n3; t += 256) {
        const double im = inv_mass[t / 3];
        if (im != 0.0) s += 0.5 * v[t] * v[t] / im;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

}  // namespace

static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

void launch_guess_features(hipStream_t s, int nw, const double* pos, const int* type, const double* anchor,
                           GuessGeom G, const int* cell_of, const int* cell_start, const int* members, float* feat,
                           float* frame) {
    if (nw > 0)
        hipLaunchKernelGGL(k_guess_features, nblk((long long)3 * nw * kGuessLanes, 256), 256, 0, s, nw, pos, type,
                           anchor, G, cell_of, cell_start, members, feat, frame);
}
void launch_guess_targets(hipStream_t s, int nw, const float* frame, const double* mu, const double* base,
                          const double* anchor, const double* ref, float* targ) {
    if (nw > 0)
        hipLaunchKernelGGL(k_guess_targets, nblk(3 * nw, 256), 256, 0, s, nw, frame, mu, base, anchor, ref, targ);
}
void launch_guess_apply(hipStream_t s, int nw, const float* frame, const float* y, float scale, const double* base,
                        const double* anchor, const double* ref, double* mu0) {
    if (nw > 0)
        hipLaunchKernelGGL(k_guess_apply, nblk(4 * nw, 256), 256, 0, s, nw, frame, y, scale, base, anchor, ref, mu0);
}
void launch_vec_sub(hipStream_t s, int n3, const double* a, const double* b, double* out) {
    if (n3 > 0) hipLaunchKernelGGL(k_vec_sub, nblk(n3, 256), 256, 0, s, n3, a, b, out);
}
int mlp_waves(int nrows) { return nblk(nrows, 32 * kWaves) * kWaves; }
void launch_mlp_infer(hipStream_t s, int nrows, const float* feat, const GuessWeights& W, float* y) {
    if (nrows > 0)
        hipLaunchKernelGGL(k_mlp<0>, nblk(nrows, 32 * kWaves), 64 * kWaves, 0, s, nrows, nullptr, feat, nullptr, W, y,
                           nullptr, nullptr);
}
void launch_mlp_grad(hipStream_t s, int nrows, const int* idx, const float* feat, const float* targ,
                     const GuessWeights& W, float* grad, float* lossp) {
    if (nrows > 0)
        hipLaunchKernelGGL(k_mlp<1>, nblk(nrows, 32 * kWaves), 64 * kWaves, 0, s, nrows, idx, feat, targ, W, nullptr,
                           grad, lossp);
}
void launch_adam(hipStream_t s, int nwaves, const float* grad, const float* lossp, float* w, float* m, float* v, int t,
                 float lr, const GuessWeights& W, float* loss_out) {
    hipLaunchKernelGGL(k_adam, nblk(kGuessParams, 256), 256, 0, s, nwaves, grad, lossp, w, m, v, t, lr, W, loss_out);
}
void launch_publish(hipStream_t s, const float* w, const GuessWeights& W) {
    hipLaunchKernelGGL(k_publish, nblk(kGuessParams, 256), 256, 0, s, w, W);
}
int colsum_blocks(size_t nrows, size_t rows_per_block) { return (int)((nrows + rows_per_block - 1) / rows_per_block); }
void launch_colsums(hipStream_t s, size_t nrows, int ncol, const float* a, size_t rows_per_block, double* out) {
    if (nrows > 0)
        hipLaunchKernelGGL(k_colsums, colsum_blocks(nrows, rows_per_block), 64 * nblk(ncol, 64), 0, s, nrows, ncol, a,
                           rows_per_block, out);
}
void launch_md_kick(hipStream_t s, int n, const double* f, const double* inv_mass, double h, double* v) {
    hipLaunchKernelGGL(k_md_kick, nblk(3 * n, 256), 256, 0, s, 3 * n, f, inv_mass, h, v);
}
void launch_md_drift(hipStream_t s, int n, const double* v, const double* inv_mass, double h, double* x) {
    hipLaunchKernelGGL(k_md_drift, nblk(3 * n, 256), 256, 0, s, 3 * n, v, inv_mass, h, x);
}
void launch_md_thermostat(hipStream_t s, int n, const double* inv_mass, double a, double kT, unsigned long long seed,
                          unsigned long long step, double* v) {
    hipLaunchKernelGGL(k_md_thermostat, nblk(3 * n, 256), 256, 0, s, 3 * n, inv_mass, a, kT, seed, step, v);
}
void launch_md_kinetic(hipStream_t s, int n, const double* v, const double* inv_mass, double* out) {
    hipLaunchKernelGGL(k_md_kinetic, 1, 256, 0, s, 3 * n, v, inv_mass, out);
}

}  // namespace mbpol
