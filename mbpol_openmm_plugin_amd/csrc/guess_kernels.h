// Types and launchers of the learned dipole guess and the Langevin dynamics
// (kernels_guess.hip).
#pragma once
#include "kernels.h"

namespace mbpol {

// |alpha E_fixed| below this (e nm) counts as this in the relative outputs
constexpr double kGuessFloor = 1e-3;
constexpr int kGuessRbf = 16, kGuessFeat = 3 * kGuessRbf, kGuessHid = 64, kGuessOut = 3;
// parameter vector: W1 [48][64], b1 [64], W2 [64][64], b2 [64], W3 [64][3], b3 [3] (matrices [in][out])
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kGuessFeat * kGuessHid, kOffW2 = kOffB1 + kGuessHid,
              kOffB2 = kOffW2 + kGuessHid * kGuessHid, kOffW3 = kOffB2 + kGuessHid,
              kOffB3 = kOffW3 + kGuessHid * kGuessOut, kGuessParams = kOffB3 + kGuessOut;

// device view of the network: BF16 operand copies of W1 (transposed) and W2
// (both orders), FP32 pointers into the master parameters for the rest, and
// the input normalisation
struct GuessWeights {
    const __bf16* w1t;  // [64][48]: w1t[j][f] = W1[f][j]
    const __bf16* w2;   // [64][64]: w2[k][j] = W2[k][j]
    const __bf16* w2t;  // [64][64]: w2t[j][k] = W2[k][j]
    const float *b1, *b2, *w3, *b3;
    const float *mean, *inv_std;  // [48]
    float inv_tscale;             // training targets are divided by the output scale
    __bf16 *w1t_out, *w2_out, *w2t_out;
};

#ifdef __HIPCC__
// write the operand copies of parameter p
__device__ __forceinline__ void guess_publish(int p, float v, const GuessWeights& W) {
    if (p < kOffB1) {
        const int f = p / kGuessHid, j = p % kGuessHid;
        W.w1t_out[j * kGuessFeat + f] = (__bf16)v;
    } else if (p >= kOffW2 && p < kOffB2) {
        const int k = (p - kOffW2) / kGuessHid, j = (p - kOffW2) % kGuessHid;
        W.w2_out[k * kGuessHid + j] = (__bf16)v;
        W.w2t_out[j * kGuessHid + k] = (__bf16)v;
    }
}
#endif

// neighbour search of the features: the electrostatics site cells when their
// grid reaches the feature cutoff, else all sites
struct GuessGeom {
    int periodic;
    d3 box, inv_box;
    int use_cells, reach;
    int3 nc;
};

// features of every polarisable atom, and its frame (e1 along the anchor vector of its site)
void launch_guess_features(hipStream_t s, int nw, const double* pos, const int* type, const double* anchor,
                           GuessGeom G, const int* cell_of, const int* cell_start, const int* members, float* feat,
                           float* frame);
// targets R (mu - base) / |anchor| and guesses base + |anchor| scale R^T y; ref (optional): the
// units |anchor| are at least 1% of |ref|
void launch_guess_targets(hipStream_t s, int nw, const float* frame, const double* mu, const double* base,
                          const double* anchor, const double* ref, float* targ);
void launch_guess_apply(hipStream_t s, int nw, const float* frame, const float* y, float scale, const double* base,
                        const double* anchor, const double* ref, double* mu0);
void launch_vec_sub(hipStream_t s, int n3, const double* a, const double* b, double* out);
int mlp_waves(int nrows);
void launch_mlp_infer(hipStream_t s, int nrows, const float* feat, const GuessWeights& W, float* y);
void launch_mlp_grad(hipStream_t s, int nrows, const int* idx, const float* feat, const float* targ,
                     const GuessWeights& W, float* grad, float* lossp);
void launch_adam(hipStream_t s, int nwaves, const float* grad, const float* lossp, float* w, float* m, float* v, int t,
                 float lr, const GuessWeights& W, float* loss_out);
void launch_publish(hipStream_t s, const float* w, const GuessWeights& W);
int colsum_blocks(size_t nrows, size_t rows_per_block);
void launch_colsums(hipStream_t s, size_t nrows, int ncol, const float* a, size_t rows_per_block, double* out);

void launch_md_kick(hipStream_t s, int n, const double* f, const double* inv_mass, double h, double* v);
void launch_md_drift(hipStream_t s, int n, const double* v, const double* inv_mass, double h, double* x);
void launch_md_thermostat(hipStream_t s, int n, const double* inv_mass, double a, double kT, unsigned long long seed,
                          unsigned long long step, double* v);
void launch_md_kinetic(hipStream_t s, int n, const double* v, const double* inv_mass, double* out);

}  // namespace mbpol
