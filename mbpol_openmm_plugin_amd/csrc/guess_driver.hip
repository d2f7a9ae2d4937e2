// Host side of the learned dipole guess and of the Langevin dynamics
// (kernels in kernels_guess.hip).  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/mbpol_hip.h"
#include "elec.hpp"
#include "guess_kernels.h"

using namespace mbpol;

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t _e = (expr);                                                                         \
        if (_e != hipSuccess) return set_error(MBPOL_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

struct mbpol_guess {
    mbpol_context* ctx = nullptr;
    int nw = 0, natoms = 0, max_snap = 0, nsnap = 0;
    // MBPOL_GUESS_KIND_ALPHA_E: a correction to alpha E_fixed, anchored on it;
    // MBPOL_GUESS_KIND_PREVIOUS: a correction to the previous evaluation's dipoles,
    // anchored on the change of alpha E_fixed since that evaluation
    int kind = MBPOL_GUESS_KIND_ALPHA_E;
    DevBuf anchor;  // the anchor vectors of the previous-step kind, per site
    DevBuf pred;    // a guess computed on request (mbpol_guess_predict_last)
    uint64_t seed = 0;
    // dataset: raw features and targets (R (mu - alpha E_fixed)) of every stored snapshot
    DevBuf feat_all, targ_all;
    // one evaluation: features, frames, network outputs
    DevBuf feat, frame, y;
    // parameters (FP32 master), Adam moments, BF16 operand copies, normalisation
    DevBuf w, m, v, w1t, w2, w2t, mean, inv_std;
    DevBuf grad, lossp, loss, idx, colsum;
    float scale = 1.0f;
    bool normalised = false;
    int adam_t = 0;

    GuessWeights view() const {
        GuessWeights W{};
        W.w1t = w1t.as<__bf16>();
        W.w2 = w2.as<__bf16>();
        W.w2t = w2t.as<__bf16>();
        const float* p = w.as<float>();
        W.b1 = p + kOffB1;
        W.b2 = p + kOffB2;
        W.w3 = p + kOffW3;
        W.b3 = p + kOffB3;
        W.mean = mean.as<float>();
        W.inv_std = inv_std.as<float>();
        W.inv_tscale = 1.0f / scale;
        W.w1t_out = w1t.as<__bf16>();
        W.w2_out = w2.as<__bf16>();
        W.w2t_out = w2t.as<__bf16>();
        return W;
    }
};

struct mbpol_md {
    mbpol_context* ctx = nullptr;
    int n = 0;
    double T = 300.0, gamma = 1.0, dt = 0.0005;
    uint64_t seed = 0, step = 0;
    uint32_t components = MBPOL_ALL;
    DevBuf inv_mass, vel, ke;
    double last_e[6] = {0, 0, 0, 0, 0, 0};
};

namespace {

constexpr double kBoltzmann = 0.0083144626;  // kJ/mol/K

// features of the positions in the context's buffer, and frames along `anchor`
// (nullptr: alpha E_fixed), on the context's stream
int compute_features(mbpol_guess* g, const double* anchor = nullptr) {
    mbpol_context* ctx = g->ctx;
    ElecView v;
    if (int e = elec_view(ctx, v)) return e;
    if (!anchor) anchor = v.efix_pol;
    GuessGeom G{};
    G.periodic = v.periodic;
    G.box = v.box;
    G.inv_box = v.inv_box;
    G.use_cells = v.use_cells;
    G.reach = v.reach;
    G.nc = v.nc;
    launch_guess_features(ctx->stream, g->nw, ctx->pos.as<double>(), v.type, anchor, G, v.cell_of, v.cell_start,
                          v.members, g->feat.as<float>(), g->frame.as<float>());
    return MBPOL_OK;
}

int sync(mbpol_context* ctx) {
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MBPOL_OK;
}

// input means and inverse deviations, and the output scale (rms target), from the stored snapshots
int set_normalisation(mbpol_guess* g, int nsnap) {
    hipStream_t st = g->ctx->stream;
    const size_t rows = (size_t)nsnap * g->natoms;
    const size_t per = 4096;
    const int nb = colsum_blocks(rows, per);
    HIP_TRY(g->colsum.ensure(sizeof(double) * 2 * kGuessFeat * (size_t)std::max(nb, 1)));
    std::vector<double> h(2 * kGuessFeat * (size_t)nb);
    std::vector<float> mean(kGuessFeat), istd(kGuessFeat);
    launch_colsums(st, rows, kGuessFeat, g->feat_all.as<float>(), per, g->colsum.as<double>());
    HIP_TRY(hipMemcpyAsync(h.data(), g->colsum.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    if (int e = sync(g->ctx)) return e;
    for (int c = 0; c < kGuessFeat; ++c) {
        double s = 0, s2 = 0;
        for (int b = 0; b < nb; ++b) {
            s += h[(size_t)(b * kGuessFeat + c) * 2];
            s2 += h[(size_t)(b * kGuessFeat + c) * 2 + 1];
        }
        const double mu = s / rows, var = std::max(s2 / rows - mu * mu, 0.0);
        mean[c] = (float)mu;
        istd[c] = (float)(1.0 / std::max(std::sqrt(var), 1e-3));
    }
    launch_colsums(st, rows, 3, g->targ_all.as<float>(), per, g->colsum.as<double>());
    HIP_TRY(hipMemcpyAsync(h.data(), g->colsum.p, sizeof(double) * 2 * 3 * nb, hipMemcpyDeviceToHost, st));
    if (int e = sync(g->ctx)) return e;
    double s2 = 0;
    for (int b = 0; b < nb; ++b)
        for (int c = 0; c < 3; ++c) s2 += h[(size_t)(b * 3 + c) * 2 + 1];
    g->scale = (float)std::max(std::sqrt(s2 / (3.0 * rows)), 1e-12);
    HIP_TRY(copy_on(g->ctx->stream, g->mean.p, mean.data(), sizeof(float) * kGuessFeat));
    HIP_TRY(copy_on(g->ctx->stream, g->inv_std.p, istd.data(), sizeof(float) * kGuessFeat));
    g->normalised = true;
    return MBPOL_OK;
}

}  // namespace

// the learned guess of one evaluation (called from the electrostatics set-up)
int mbpol::guess_predict(mbpol_context* ctx, double* mu0) {
    mbpol_guess* g = ctx->guess;
    if (!g) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "no learned guess attached");
    ElecView v;
    if (int e = elec_view(ctx, v)) return e;
    hipStream_t st = ctx->stream;
    if (g->kind == MBPOL_GUESS_KIND_PREVIOUS) {
        if (!v.have_previous) {  // no previous evaluation yet: the reference's start
            HIP_TRY(hipMemcpyAsync(mu0, v.efix_pol, sizeof(double) * 3 * v.n, hipMemcpyDeviceToDevice, st));
            return MBPOL_OK;
        }
        launch_vec_sub(st, 3 * v.n, v.efix_pol, v.efix_prev, g->anchor.as<double>());
        if (int e = compute_features(g, g->anchor.as<double>())) return e;
        launch_mlp_infer(st, g->natoms, g->feat.as<float>(), g->view(), g->y.as<float>());
        launch_guess_apply(st, g->nw, g->frame.as<float>(), g->y.as<float>(), g->scale, v.mu_prev,
                           g->anchor.as<double>(), v.efix_pol, mu0);
        return MBPOL_OK;
    }
    if (int e = compute_features(g)) return e;
    launch_mlp_infer(st, g->natoms, g->feat.as<float>(), g->view(), g->y.as<float>());
    launch_guess_apply(st, g->nw, g->frame.as<float>(), g->y.as<float>(), g->scale, v.efix_pol, v.efix_pol, nullptr,
                       mu0);
    return MBPOL_OK;
}

extern "C" {

int mbpol_guess_create(mbpol_context* ctx, int32_t max_snapshots, uint64_t seed, mbpol_guess** out) {
    return mbpol_guess_create_kind(ctx, max_snapshots, seed, MBPOL_GUESS_KIND_ALPHA_E, out);
}

int mbpol_guess_create_kind(mbpol_context* ctx, int32_t max_snapshots, uint64_t seed, int32_t kind,
                            mbpol_guess** out) {
    if (!ctx || !out || max_snapshots < 0) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "bad argument");
    if (kind != MBPOL_GUESS_KIND_ALPHA_E && kind != MBPOL_GUESS_KIND_PREVIOUS)
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "unknown guess kind %d", kind);
    ElecView v;
    if (int e = elec_view(ctx, v)) return e;
    if (!v.water_layout)
        return set_error(MBPOL_ERR_INVALID_ARGUMENT,
                         "the learned guess needs water sites in O H H M order with charge redistribution");
    HIP_TRY(hipSetDevice(ctx->device));
    mbpol_guess* g = new mbpol_guess();
    g->ctx = ctx;
    g->nw = v.nw;
    g->natoms = 3 * v.nw;
    g->max_snap = max_snapshots;
    g->seed = seed;
    g->kind = kind;
    auto fail = [&](int code) {
        delete g;
        return code;
    };
    const size_t na = (size_t)std::max(g->natoms, 1);
    if (g->feat_all.ensure(sizeof(float) * kGuessFeat * na * std::max(max_snapshots, 1)) ||
        g->targ_all.ensure(sizeof(float) * 3 * na * std::max(max_snapshots, 1)) ||
        g->feat.ensure(sizeof(float) * kGuessFeat * na) || g->frame.ensure(sizeof(float) * 9 * na) ||
        g->y.ensure(sizeof(float) * 3 * na) || g->w.ensure(sizeof(float) * kGuessParams) ||
        g->m.ensure(sizeof(float) * kGuessParams) || g->v.ensure(sizeof(float) * kGuessParams) ||
        g->w1t.ensure(2 * kGuessFeat * kGuessHid) || g->w2.ensure(2 * kGuessHid * kGuessHid) ||
        g->w2t.ensure(2 * kGuessHid * kGuessHid) || g->mean.ensure(sizeof(float) * kGuessFeat) ||
        g->inv_std.ensure(sizeof(float) * kGuessFeat) || g->loss.ensure(sizeof(float) * 4) ||
        g->anchor.ensure(sizeof(double) * 3 * std::max(v.n, 1)) || g->pred.ensure(sizeof(double) * 3 * std::max(v.n, 1)))
        return fail(set_error(MBPOL_ERR_HIP, "device allocation for the learned guess failed"));
    // the previous-step kind reads the electrostatics' history of dipoles and fields
    if (kind == MBPOL_GUESS_KIND_PREVIOUS) elec_keep_history(ctx, true);
    // He-style initial weights (deviation 1/sqrt(fan_in)), zero biases, a small last layer:
    // the untrained network proposes alpha E_fixed plus a small correction
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<float> w(kGuessParams, 0.f);
    for (int p = kOffW1; p < kOffB1; ++p) w[p] = (float)(nd(rng) / std::sqrt((double)kGuessFeat));
    for (int p = kOffW2; p < kOffB2; ++p) w[p] = (float)(nd(rng) / std::sqrt((double)kGuessHid));
    for (int p = kOffW3; p < kOffB3; ++p) w[p] = (float)(0.1 * nd(rng) / std::sqrt((double)kGuessHid));
    std::vector<float> zero(kGuessParams, 0.f), mean(kGuessFeat, 0.f), one(kGuessFeat, 1.f);
    if (copy_on(ctx->stream, g->w.p, w.data(), sizeof(float) * kGuessParams) ||
        copy_on(ctx->stream, g->m.p, zero.data(), sizeof(float) * kGuessParams) ||
        copy_on(ctx->stream, g->v.p, zero.data(), sizeof(float) * kGuessParams) ||
        copy_on(ctx->stream, g->mean.p, mean.data(), sizeof(float) * kGuessFeat) ||
        copy_on(ctx->stream, g->inv_std.p, one.data(), sizeof(float) * kGuessFeat))
        return fail(set_error(MBPOL_ERR_HIP, "hipMemcpy failed"));
    launch_publish(ctx->stream, g->w.as<float>(), g->view());
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(set_error(MBPOL_ERR_HIP, "publish failed"));
    *out = g;
    return MBPOL_OK;
}

void mbpol_guess_destroy(mbpol_guess* g) {
    if (!g) return;
    if (g->ctx && g->ctx->guess == g) {
        g->ctx->guess = nullptr;
        if (g->ctx->elec) (void)mbpol_context_set_dipole_guess(g->ctx, MBPOL_GUESS_ALPHA_E, nullptr);
    }
    if (g->ctx && g->kind == MBPOL_GUESS_KIND_PREVIOUS) elec_keep_history(g->ctx, false);
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    delete g;
}

int mbpol_guess_add_snapshot(mbpol_guess* g) {
    if (!g) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null guess");
    if (g->nsnap >= g->max_snap) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "snapshot store is full (%d)", g->max_snap);
    mbpol_context* ctx = g->ctx;
    if (!ctx->stats.converged) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "the last evaluate has no converged dipoles");
    HIP_TRY(hipSetDevice(ctx->device));
    ElecView v;
    if (int e = elec_view(ctx, v)) return e;
    const size_t off = (size_t)g->nsnap * g->natoms;
    if (g->kind == MBPOL_GUESS_KIND_PREVIOUS) {
        // the pair (previous evaluation, this one): base = the previous converged dipoles,
        // anchor = the change of alpha E_fixed between the two
        if (!v.have_previous2)
            return set_error(MBPOL_ERR_INVALID_ARGUMENT, "a snapshot of this kind needs two earlier evaluations");
        launch_vec_sub(ctx->stream, 3 * v.n, v.efix_pol, v.efix_prev2, g->anchor.as<double>());
        if (int e = compute_features(g, g->anchor.as<double>())) return e;
        launch_guess_targets(ctx->stream, g->nw, g->frame.as<float>(), v.mu, v.mu_prev2, g->anchor.as<double>(),
                             v.efix_pol, g->targ_all.as<float>() + off * 3);
    } else {
        if (int e = compute_features(g)) return e;
        launch_guess_targets(ctx->stream, g->nw, g->frame.as<float>(), v.mu, v.efix_pol, v.efix_pol, nullptr,
                             g->targ_all.as<float>() + off * 3);
    }
    HIP_TRY(hipMemcpyAsync(g->feat_all.as<float>() + off * kGuessFeat, g->feat.p, sizeof(float) * kGuessFeat * g->natoms,
                           hipMemcpyDeviceToDevice, ctx->stream));
    if (int e = sync(ctx)) return e;
    ++g->nsnap;
    return MBPOL_OK;
}

int mbpol_guess_predict_last(mbpol_guess* g, double* dipoles) {
    if (!g || !dipoles) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "bad argument");
    if (!g->normalised) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "the network has not been trained");
    mbpol_context* ctx = g->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    ElecView v;
    if (int e = elec_view(ctx, v)) return e;
    hipStream_t st = ctx->stream;
    double* out = g->pred.as<double>();
    if (g->kind == MBPOL_GUESS_KIND_PREVIOUS) {
        if (!v.have_previous2)
            return set_error(MBPOL_ERR_INVALID_ARGUMENT, "a guess of this kind needs two earlier evaluations");
        launch_vec_sub(st, 3 * v.n, v.efix_pol, v.efix_prev2, g->anchor.as<double>());
        if (int e = compute_features(g, g->anchor.as<double>())) return e;
        launch_mlp_infer(st, g->natoms, g->feat.as<float>(), g->view(), g->y.as<float>());
        launch_guess_apply(st, g->nw, g->frame.as<float>(), g->y.as<float>(), g->scale, v.mu_prev2,
                           g->anchor.as<double>(), v.efix_pol, out);
    } else {
        if (int e = compute_features(g)) return e;
        launch_mlp_infer(st, g->natoms, g->feat.as<float>(), g->view(), g->y.as<float>());
        launch_guess_apply(st, g->nw, g->frame.as<float>(), g->y.as<float>(), g->scale, v.efix_pol, v.efix_pol,
                           nullptr, out);
    }
    if (int e = sync(ctx)) return e;
    HIP_TRY(copy_on(g->ctx->stream, dipoles, out, sizeof(double) * 3 * v.n));
    return MBPOL_OK;
}

int mbpol_guess_num_snapshots(const mbpol_guess* g, int32_t* count) {
    if (!g || !count) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null argument");
    *count = g->nsnap;
    return MBPOL_OK;
}

int mbpol_guess_get_snapshot(const mbpol_guess* g, int32_t k, float* features, float* targets) {
    if (!g || k < 0 || k >= g->nsnap) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "no snapshot %d", k);
    HIP_TRY(hipSetDevice(g->ctx->device));
    HIP_TRY(hipStreamSynchronize(g->ctx->stream));
    const size_t off = (size_t)k * g->natoms;
    if (features)
        HIP_TRY(copy_on(g->ctx->stream, features, g->feat_all.as<float>() + off * kGuessFeat, sizeof(float) * kGuessFeat * g->natoms));
    if (targets)
        HIP_TRY(copy_on(g->ctx->stream, targets, g->targ_all.as<float>() + off * 3, sizeof(float) * 3 * g->natoms));
    return MBPOL_OK;
}

int mbpol_guess_train(mbpol_guess* g, int32_t epochs, int32_t batch, double learning_rate, float* loss_per_epoch) {
    if (!g) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null guess");
    return mbpol_guess_train_range(g, g->nsnap, epochs, batch, learning_rate, loss_per_epoch);
}

int mbpol_guess_train_range(mbpol_guess* g, int32_t ntrain, int32_t epochs, int32_t batch, double learning_rate,
                            float* loss_per_epoch) {
    if (!g || epochs < 0 || batch < 32) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "bad argument");
    if (ntrain <= 0 || ntrain > g->nsnap) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "no snapshots to train on");
    mbpol_context* ctx = g->ctx;
    hipStream_t st = ctx->stream;
    HIP_TRY(hipSetDevice(ctx->device));
    if (!g->normalised)
        if (int e = set_normalisation(g, ntrain)) return e;
    const size_t rows = (size_t)ntrain * g->natoms;
    if (rows > (size_t)INT32_MAX) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "too many samples");
    const int waves = mlp_waves(batch);
    HIP_TRY(g->grad.ensure(sizeof(float) * kGuessParams * (size_t)waves));
    HIP_TRY(g->lossp.ensure(sizeof(float) * waves));
    HIP_TRY(g->idx.ensure(sizeof(int) * rows));
    const size_t steps = rows / (size_t)batch;  // whole batches; the remainder is reshuffled into the next epoch
    HIP_TRY(g->loss.ensure(sizeof(float) * std::max<size_t>(steps, 1)));
    std::mt19937_64 rng(g->seed + 0x5bd1e995ull * (uint64_t)(g->adam_t + 1));
    std::vector<int> perm(rows);
    std::vector<float> hl(std::max<size_t>(steps, 1));
    const GuessWeights W = g->view();
    for (int ep = 0; ep < epochs; ++ep) {
        std::iota(perm.begin(), perm.end(), 0);
        std::shuffle(perm.begin(), perm.end(), rng);
        HIP_TRY(hipMemcpyAsync(g->idx.p, perm.data(), sizeof(int) * rows, hipMemcpyHostToDevice, st));
        for (size_t s = 0; s < steps; ++s) {
            launch_mlp_grad(st, batch, g->idx.as<int>() + s * batch, g->feat_all.as<float>(), g->targ_all.as<float>(), W,
                            g->grad.as<float>(), g->lossp.as<float>());
            launch_adam(st, waves, g->grad.as<float>(), g->lossp.as<float>(), g->w.as<float>(), g->m.as<float>(),
                        g->v.as<float>(), ++g->adam_t, (float)learning_rate, W, g->loss.as<float>() + s);
        }
        HIP_TRY(hipMemcpyAsync(hl.data(), g->loss.p, sizeof(float) * steps, hipMemcpyDeviceToHost, st));
        if (int e = sync(ctx)) return e;
        double tot = 0;
        for (size_t s = 0; s < steps; ++s) tot += hl[s];
        if (loss_per_epoch) loss_per_epoch[ep] = (float)(steps ? tot / steps : 0.0);
    }
    return MBPOL_OK;
}

int mbpol_guess_validation_loss(mbpol_guess* g, int32_t first, int32_t count, int32_t batch, double* loss) {
    if (!g || !loss || first < 0 || count <= 0 || first + count > g->nsnap || batch < 32)
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "bad argument");
    if (!g->normalised) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "the network has not been trained");
    mbpol_context* ctx = g->ctx;
    hipStream_t st = ctx->stream;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t r0 = (size_t)first * g->natoms, rows = (size_t)count * g->natoms;
    if (r0 + rows > (size_t)INT32_MAX) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "too many samples");
    const int waves = mlp_waves(batch);
    HIP_TRY(g->grad.ensure(sizeof(float) * kGuessParams * (size_t)waves));
    HIP_TRY(g->lossp.ensure(sizeof(float) * waves));
    HIP_TRY(g->idx.ensure(sizeof(int) * rows));
    std::vector<int> id(rows);
    std::iota(id.begin(), id.end(), (int)r0);
    HIP_TRY(hipMemcpyAsync(g->idx.p, id.data(), sizeof(int) * rows, hipMemcpyHostToDevice, st));
    std::vector<float> hl(waves);
    double tot = 0;
    for (size_t b = 0; b < rows; b += (size_t)batch) {
        const int nb = (int)std::min<size_t>(batch, rows - b);
        launch_mlp_grad(st, nb, g->idx.as<int>() + b, g->feat_all.as<float>(), g->targ_all.as<float>(), g->view(),
                        g->grad.as<float>(), g->lossp.as<float>());
        HIP_TRY(hipMemcpyAsync(hl.data(), g->lossp.p, sizeof(float) * mlp_waves(nb), hipMemcpyDeviceToHost, st));
        if (int e = sync(ctx)) return e;
        double s = 0;
        for (int k = 0; k < mlp_waves(nb); ++k) s += hl[k];
        tot += s * nb;  // per-batch mean -> sum
    }
    *loss = tot / (double)rows;
    return MBPOL_OK;
}

int mbpol_guess_attach(mbpol_guess* g, int32_t on) {
    if (!g) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null guess");
    mbpol_context* ctx = g->ctx;
    if (on) {
        ctx->guess = g;
        return mbpol_context_set_dipole_guess(ctx, MBPOL_GUESS_LEARNED, nullptr);
    }
    if (ctx->guess == g) ctx->guess = nullptr;
    return mbpol_context_set_dipole_guess(ctx, MBPOL_GUESS_ALPHA_E, nullptr);
}

int mbpol_guess_get_params(const mbpol_guess* g, float* params, float* norm) {
    if (!g) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null guess");
    HIP_TRY(hipSetDevice(g->ctx->device));
    HIP_TRY(hipStreamSynchronize(g->ctx->stream));
    if (params) HIP_TRY(copy_on(g->ctx->stream, params, g->w.p, sizeof(float) * kGuessParams));
    if (norm) {
        HIP_TRY(copy_on(g->ctx->stream, norm, g->mean.p, sizeof(float) * kGuessFeat));
        HIP_TRY(copy_on(g->ctx->stream, norm + kGuessFeat, g->inv_std.p, sizeof(float) * kGuessFeat));
        norm[2 * kGuessFeat] = g->scale;
    }
    return MBPOL_OK;
}

int mbpol_guess_set_params(mbpol_guess* g, const float* params, const float* norm) {
    if (!g) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null guess");
    HIP_TRY(hipSetDevice(g->ctx->device));
    HIP_TRY(hipStreamSynchronize(g->ctx->stream));
    if (params) HIP_TRY(copy_on(g->ctx->stream, g->w.p, params, sizeof(float) * kGuessParams));
    if (norm) {
        if (!(norm[2 * kGuessFeat] > 0.f)) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "output scale must be > 0");
        HIP_TRY(copy_on(g->ctx->stream, g->mean.p, norm, sizeof(float) * kGuessFeat));
        HIP_TRY(copy_on(g->ctx->stream, g->inv_std.p, norm + kGuessFeat, sizeof(float) * kGuessFeat));
        g->scale = norm[2 * kGuessFeat];
        g->normalised = true;
    }
    launch_publish(g->ctx->stream, g->w.as<float>(), g->view());
    return sync(g->ctx);
}

int mbpol_guess_gradient(mbpol_guess* g, int32_t n, const float* x, const float* t, float* grad, float* loss) {
    if (!g || n <= 0 || !x || !t || !grad) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "bad argument");
    hipStream_t st = g->ctx->stream;
    HIP_TRY(hipSetDevice(g->ctx->device));
    DevBuf dx, dt, di, dg, dl;
    const int waves = mlp_waves(n);
    HIP_TRY(dx.ensure(sizeof(float) * kGuessFeat * n));
    HIP_TRY(dt.ensure(sizeof(float) * 3 * n));
    HIP_TRY(di.ensure(sizeof(int) * n));
    HIP_TRY(dg.ensure(sizeof(float) * kGuessParams * waves));
    HIP_TRY(dl.ensure(sizeof(float) * waves));
    std::vector<int> id(n);
    std::iota(id.begin(), id.end(), 0);
    HIP_TRY(copy_on(g->ctx->stream, dx.p, x, sizeof(float) * kGuessFeat * n));
    HIP_TRY(copy_on(g->ctx->stream, dt.p, t, sizeof(float) * 3 * n));
    HIP_TRY(copy_on(g->ctx->stream, di.p, id.data(), sizeof(int) * n));
    launch_mlp_grad(st, n, di.as<int>(), dx.as<float>(), dt.as<float>(), g->view(), dg.as<float>(), dl.as<float>());
    std::vector<float> hg((size_t)kGuessParams * waves), hl(waves);
    HIP_TRY(hipMemcpyAsync(hg.data(), dg.p, sizeof(float) * hg.size(), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(hl.data(), dl.p, sizeof(float) * waves, hipMemcpyDeviceToHost, st));
    if (int e = sync(g->ctx)) return e;
    for (int p = 0; p < kGuessParams; ++p) {
        double s = 0;
        for (int k = 0; k < waves; ++k) s += hg[(size_t)k * kGuessParams + p];
        grad[p] = (float)s;
    }
    if (loss) {
        double s = 0;
        for (int k = 0; k < waves; ++k) s += hl[k];
        *loss = (float)s;
    }
    return MBPOL_OK;
}

int mbpol_guess_forward(mbpol_guess* g, int32_t n, const float* x, float* y) {
    if (!g || n <= 0 || !x || !y) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "bad argument");
    HIP_TRY(hipSetDevice(g->ctx->device));
    DevBuf dx, dy;
    HIP_TRY(dx.ensure(sizeof(float) * kGuessFeat * n));
    HIP_TRY(dy.ensure(sizeof(float) * 3 * n));
    HIP_TRY(copy_on(g->ctx->stream, dx.p, x, sizeof(float) * kGuessFeat * n));
    launch_mlp_infer(g->ctx->stream, n, dx.as<float>(), g->view(), dy.as<float>());
    if (int e = sync(g->ctx)) return e;
    HIP_TRY(copy_on(g->ctx->stream, y, dy.p, sizeof(float) * 3 * n));
    return MBPOL_OK;
}

int mbpol_guess_features(mbpol_guess* g, float* features, float* frames) {
    if (!g) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null guess");
    HIP_TRY(hipSetDevice(g->ctx->device));
    if (int e = compute_features(g)) return e;
    if (int e = sync(g->ctx)) return e;
    if (features) HIP_TRY(copy_on(g->ctx->stream, features, g->feat.p, sizeof(float) * kGuessFeat * g->natoms));
    if (frames) HIP_TRY(copy_on(g->ctx->stream, frames, g->frame.p, sizeof(float) * 9 * g->natoms));
    return MBPOL_OK;
}

// ---------------------------------------------------------------------------
// Langevin dynamics
// ---------------------------------------------------------------------------
int mbpol_md_create(mbpol_context* ctx, const double* masses, double temperature_K, double friction_per_ps,
                    double dt_ps, uint64_t seed, uint32_t components, mbpol_md** out) {
    if (!ctx || !masses || !out || !(dt_ps > 0) || temperature_K < 0 || friction_per_ps < 0)
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "bad argument");
    if (!ctx->pos_resident.p) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "upload positions before creating dynamics");
    HIP_TRY(hipSetDevice(ctx->device));
    const int n = ctx->sys.n;
    mbpol_md* md = new mbpol_md();
    md->ctx = ctx;
    md->n = n;
    md->T = temperature_K;
    md->gamma = friction_per_ps;
    md->dt = dt_ps;
    md->seed = seed;
    md->components = components;
    std::vector<double> im(n), v(3 * (size_t)n, 0.0);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    const double kT = kBoltzmann * temperature_K;
    for (int i = 0; i < n; ++i) {
        im[i] = masses[i] > 0 ? 1.0 / masses[i] : 0.0;
        if (im[i] != 0.0)
            for (int d = 0; d < 3; ++d) v[3 * (size_t)i + d] = std::sqrt(kT * im[i]) * nd(rng);
    }
    auto fail = [&](int code) {
        delete md;
        return code;
    };
    if (md->inv_mass.ensure(sizeof(double) * std::max(n, 1)) || md->vel.ensure(sizeof(double) * 3 * std::max(n, 1)) ||
        md->ke.ensure(sizeof(double)))
        return fail(set_error(MBPOL_ERR_HIP, "device allocation for the dynamics failed"));
    if (copy_on(md->ctx->stream, md->inv_mass.p, im.data(), sizeof(double) * n) ||
        copy_on(md->ctx->stream, md->vel.p, v.data(), sizeof(double) * 3 * n))
        return fail(set_error(MBPOL_ERR_HIP, "hipMemcpy failed"));
    if (int e = mbpol_context_evaluate_resident(ctx, md->last_e, components)) return fail(e);
    *out = md;
    return MBPOL_OK;
}

void mbpol_md_destroy(mbpol_md* md) { delete md; }

int mbpol_md_step(mbpol_md* md, int32_t nsteps, double* energies_out, double* kinetic_out) {
    if (!md || nsteps < 0) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "bad argument");
    mbpol_context* ctx = md->ctx;
    hipStream_t st = ctx->stream;
    HIP_TRY(hipSetDevice(ctx->device));
    const int n = md->n;
    double* x = ctx->pos_resident.as<double>();
    const double* f = ctx->force.as<double>();
    const double* im = md->inv_mass.as<double>();
    double* v = md->vel.as<double>();
    const double h = 0.5 * md->dt, a = std::exp(-md->gamma * md->dt), kT = kBoltzmann * md->T;
    for (int s = 0; s < nsteps; ++s) {
        launch_md_kick(st, n, f, im, h, v);
        launch_md_drift(st, n, v, im, h, x);
        launch_md_thermostat(st, n, im, a, kT, md->seed, md->step++, v);
        launch_md_drift(st, n, v, im, h, x);
        if (int e = mbpol_context_evaluate_resident(ctx, md->last_e, md->components)) return e;
        launch_md_kick(st, n, f, im, h, v);
    }
    if (energies_out) std::memcpy(energies_out, md->last_e, sizeof(md->last_e));
    if (kinetic_out) {
        launch_md_kinetic(st, n, v, im, md->ke.as<double>());
        HIP_TRY(hipMemcpyAsync(kinetic_out, md->ke.p, sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    } else {
        HIP_TRY(hipStreamSynchronize(st));
    }
    return MBPOL_OK;
}

int mbpol_md_get_positions(const mbpol_md* md, double* positions) {
    if (!md || !positions) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipSetDevice(md->ctx->device));
    HIP_TRY(hipStreamSynchronize(md->ctx->stream));
    HIP_TRY(copy_on(md->ctx->stream, positions, md->ctx->pos_resident.p, sizeof(double) * 3 * md->n));
    return MBPOL_OK;
}

}  // extern "C"
