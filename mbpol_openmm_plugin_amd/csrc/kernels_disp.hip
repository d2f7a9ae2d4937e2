// Damped-C6 dispersion pair kernel (python/mbpol.xml:37-108):
//   E = -C6 tt6 / r^6,  tt6 = 1 - exp(-d6 r) sum_{k<=6} (d6 r)^k / k!
// One wave per active particle; lanes stride over its candidates (27 cells
// or every particle), queue those within the cutoff and evaluate them in full
// waves; the force on the wave's particle is reduced in registers and written
// once, each pair is visited from both ends and its energy halved.
// Arithmetic type T: float (mixed) or double.
#include "device_common.h"
#include "disp_kernels.h"

namespace mbpol {

__global__ void k_disp_cells(int na, const int* __restrict__ act, const double* __restrict__ pos, d3 L, int3 nc,
                             int* __restrict__ cell_of, int* __restrict__ cell_count) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= na) return;
    const int i = act[a];
    const double p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    const double Ld[3] = {L.x, L.y, L.z};
    const int ncd[3] = {nc.x, nc.y, nc.z};
    int c[3];
    for (int d = 0; d < 3; ++d) {
        const double q = p[d] - Ld[d] * floor(p[d] / Ld[d]);
        const int v = (int)floor(q * ncd[d] / Ld[d]);
        c[d] = v < 0 ? 0 : (v >= ncd[d] ? ncd[d] - 1 : v);
    }
    const int cell = (c[0] * nc.y + c[1]) * nc.z + c[2];
    cell_of[a] = cell;
    atomicAdd(&cell_count[cell], 1);
}

__global__ void k_disp_cell_fill(int na, const int* __restrict__ cell_of, const int* __restrict__ cell_start,
                                 int* __restrict__ cell_fill, int* __restrict__ members) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= na) return;
    const int c = cell_of[a];
    members[cell_start[c] + atomicAdd(&cell_fill[c], 1)] = a;
}

// positions and molecule * 16 + type of the active atoms in cell order (members[t] indexes act)
__global__ void k_disp_pack(int na, const int* __restrict__ members, const int* __restrict__ act,
                            const double* __restrict__ pos, const int* __restrict__ type, const int* __restrict__ mol,
                            double4* __restrict__ cpos) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= na) return;
    const int j = act[members[t]];
    cpos[t] = make_double4(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], (double)(mol[j] * 16 + type[j]));
}

template <typename T>
__global__ void __launch_bounds__(64) k_disp(DispGeom g, const int* __restrict__ act, const double* __restrict__ pos,
                                             const int* __restrict__ type, const int* __restrict__ mol,
                                             const double* __restrict__ c6, const double* __restrict__ d6,
                                             const int* __restrict__ cell_of, const int* __restrict__ cell_start,
                                             const int* __restrict__ members, const double4* __restrict__ cpos,
                                             u64* __restrict__ force, double* __restrict__ eblk,
                                             unsigned long long* __restrict__ cnt_out) {
    // The candidates within the cutoff (about one in nine of the 27-cell stencil)
    // are queued in LDS and their damped-C6 terms evaluated 64 at a time, so that
    // the exponential and the divisions run on full waves, not on the few lanes
    // of each candidate batch that found a partner.
    constexpr int kQ = 128;  // ring of queued pairs; at most 63 + 64 are pending
    __shared__ double qx[kQ], qy[kQ], qz[kQ];
    __shared__ int qt[kQ];
    const int lane = threadIdx.x;
    const unsigned long long below = (1ull << lane) - 1ull;
    const double rc2 = g.cutoff * g.cutoff;
    // minimum image by products with 1/L: the image index differs from the one
    // the quotient gives only where |dx| is within rounding of L/2, beyond the cutoff
    const double ilx = 1.0 / g.L.x, ily = 1.0 / g.L.y, ilz = 1.0 / g.L.z;
    // one atom per wave; a capped grid strides over the atoms
    for (int a = blockIdx.x; a < g.na; a += gridDim.x) {
        const int i = act[a];
        const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
        const int ti = type[i], mi = mol[i];
        double fx = 0, fy = 0, fz = 0, en = 0;
        unsigned long long cnt = 0;
        int head = 0, tail = 0;  // wave-uniform queue positions
        // the damped-C6 term of a queued pair (separation dx, dy, dz; partner type tj)
        auto pair = [&](double dx, double dy, double dz, int tj) {
            const double r2d = dx * dx + dy * dy + dz * dz;
            const T C = (T)c6[ti + g.ntypes * tj], D = (T)d6[ti + g.ntypes * tj];
            const T r2 = (T)r2d, r = sqrt(r2), x = D * r;
            const T s = (T)1 + ((T)1 + ((T)1 + ((T)1 + ((T)1 + ((T)1 + x / (T)6) * x / (T)5) * x / (T)4) * x / (T)3) * x / (T)2) * x;
            const T ex = exp_t<T>(-x);
            const T tt6 = (T)1 - s * ex;
            const T r6 = r2 * r2 * r2;
            const T x3 = x * x * x;
            const T dEdr = (T)6 * C * tt6 / (r6 * r) - C * D * (x3 * x3) * ex / ((T)720 * r6);
            const double fr = -(double)(dEdr / r);
            fx += fr * dx;
            fy += fr * dy;
            fz += fr * dz;
            en += -(double)(C * tt6 / r6);
            ++cnt;
        };
        // evaluate queued pairs 64 at a time (all of them when `all`)
        auto drain = [&](bool all) {
            while (tail - head >= 64 || (all && tail > head)) {
                __syncthreads();  // the queue entries are written
                const int k = head + lane;
                if (k < tail) {
                    const int q = k & (kQ - 1);
                    pair(qx[q], qy[q], qz[q], qt[q]);
                }
                head += min(64, tail - head);
                __syncthreads();  // read before the slots are written again
            }
        };
        // candidate q (position, molecule * 16 + type); `valid` false for a lane past the range
        auto offer = [&](bool valid, const double4& q) {
            const int key = (int)q.w, mj = key >> 4;
            double dx = xi - q.x, dy = yi - q.y, dz = zi - q.z;
            if (g.periodic) {
                dx -= floor(dx * ilx + 0.5) * g.L.x;
                dy -= floor(dy * ily + 0.5) * g.L.y;
                dz -= floor(dz * ilz + 0.5) * g.L.z;
            }
            const double r2d = dx * dx + dy * dy + dz * dz;
            // another molecule (the atom itself excluded), within the cutoff unless there is none
            const bool ok = valid && mj != mi && (g.method == 0 || r2d < rc2);
            const unsigned long long m = __ballot(ok);
            if (ok) {
                const int q = (tail + __popcll(m & below)) & (kQ - 1);
                qx[q] = dx;
                qy[q] = dy;
                qz[q] = dz;
                qt[q] = key & 15;
            }
            tail += __popcll(m);
            drain(false);
        };
        if (!g.use_cells) {
            const int b0 = g.gs ? g.gs[a] : 0, b1 = g.gs ? g.ge[a] : g.na;
            for (int b = b0; b < b1; b += 64) {
                const bool valid = b + lane < b1;
                const int j = act[valid ? b + lane : b0];
                offer(valid, make_double4(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], (double)(mol[j] * 16 + type[j])));
            }
        } else {
            const int cell = cell_of[a];
            const int cz = cell % g.nc.z, cy = (cell / g.nc.z) % g.nc.y, cx = cell / (g.nc.z * g.nc.y);
            // the three cells along z of each (X, Y) column are consecutive in the cell order:
            // one or two contiguous member ranges
            auto scan = [&](int base, int za, int zb) {
                const int b = cell_start[base + za], e = cell_start[base + zb + 1];
                for (int t = b; t < e; t += 64) {
                    const bool valid = t + lane < e;
                    offer(valid, cpos[valid ? t + lane : b]);
                }
            };
            for (int ax = -1; ax <= 1; ++ax)
                for (int ay = -1; ay <= 1; ++ay) {
                    int X = cx + ax, Y = cy + ay;
                    X = X < 0 ? X + g.nc.x : (X >= g.nc.x ? X - g.nc.x : X);
                    Y = Y < 0 ? Y + g.nc.y : (Y >= g.nc.y ? Y - g.nc.y : Y);
                    const int base = (X * g.nc.y + Y) * g.nc.z, z0 = cz - 1, z1 = cz + 1;
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
        drain(true);
        fx = wave_sum(fx);
        fy = wave_sum(fy);
        fz = wave_sum(fz);
        en = wave_sum(en);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
        if (lane == 0) {
            add_force(force, i, fx, fy, fz);
            eblk[a] = 0.5 * en;
            cnt_out[a] = cnt;
        }
    }
}

static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

void launch_disp_cells(hipStream_t s, int na, const int* act, const double* pos, d3 L, int3 nc, int* cell_of,
                       int* cell_count) {
    hipLaunchKernelGGL(k_disp_cells, nblk(na, 256), 256, 0, s, na, act, pos, L, nc, cell_of, cell_count);
}
void launch_disp_cell_fill(hipStream_t s, int na, const int* cell_of, const int* cell_start, int* cell_fill, int* members) {
    hipLaunchKernelGGL(k_disp_cell_fill, nblk(na, 256), 256, 0, s, na, cell_of, cell_start, cell_fill, members);
}
// one block: total of the per-atom pair counts (each pair counted from both atoms)
__global__ void k_count_total(int na, const unsigned long long* __restrict__ cnt, unsigned long long* __restrict__ out) {
    __shared__ unsigned long long red[256];
    unsigned long long s = 0;
    for (int a = threadIdx.x; a < na; a += 256) s += cnt[a];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

void launch_count_total(hipStream_t s, int na, const unsigned long long* cnt, unsigned long long* out) {
    hipLaunchKernelGGL(k_count_total, 1, 256, 0, s, na, cnt, out);
}
void launch_disp_pack(hipStream_t s, int na, const int* members, const int* act, const double* pos, const int* type,
                      const int* mol, double4* cpos) {
    if (na > 0) hipLaunchKernelGGL(k_disp_pack, nblk(na, 256), 256, 0, s, na, members, act, pos, type, mol, cpos);
}
void launch_disp(hipStream_t s, int precision_double, DispGeom g, const int* act, const double* pos, const int* type,
                 const int* mol, const double* c6, const double* d6, const int* cell_of, const int* cell_start,
                 const int* members, const double4* cpos, u64* force, double* eblk, unsigned long long* cnt,
                 int max_blocks) {
    if (g.na <= 0) return;
    const int grid = max_blocks > 0 && g.na > max_blocks ? max_blocks : g.na;
    if (precision_double)
        hipLaunchKernelGGL(k_disp<double>, grid, 64, 0, s, g, act, pos, type, mol, c6, d6, cell_of, cell_start, members,
                           cpos, force, eblk, cnt);
    else
        hipLaunchKernelGGL(k_disp<float>, grid, 64, 0, s, g, act, pos, type, mol, c6, d6, cell_of, cell_start, members,
                           cpos, force, eblk, cnt);
}

}  // namespace mbpol
