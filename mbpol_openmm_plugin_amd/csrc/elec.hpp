// Declarations shared between the evaluation driver (mbpol_api.hip) and the
// electrostatics / dispersion drivers.
#pragma once
#include <functional>
#include "context.hpp"

namespace mbpol {

constexpr double kFixInvHost = 1.0 / 4294967296.0;

int set_error(int code, const char* fmt, ...);
// exclusive scans on stream s (nullptr: the context's main stream)
int scan_i32(mbpol_context* ctx, const int* in, int* out, int n, hipStream_t s = nullptr);
int scan_i64(mbpol_context* ctx, const long long* in, long long* out, int n, hipStream_t s = nullptr);
void timer_start(mbpol_context* ctx, int cat);
void timer_stop(mbpol_context* ctx, int cat);
// events for one timed kernel launch (null pair when timing is off) and their bookkeeping after the launch
KEvents timer_kernel(mbpol_context* ctx, int cat);
void timer_kernel_done(mbpol_context* ctx, int cat, const KEvents& ev);

// electrostatics (elec_driver.hip)
int elec_create(mbpol_context* ctx);
void elec_destroy(mbpol_context* ctx);
int elec_set_box(mbpol_context* ctx);
// after_setup (optional) runs once the set-up phase is enqueued, before the dipole solve.
// fixed_steps > 0: a conjugate-gradient solve of that many field evaluations with no
// host read (the form a HIP graph captures); elec_graph_finish reads its outcome.
int elec_evaluate(mbpol_context* ctx, u64* energy, const std::function<int()>& after_setup = nullptr,
                  int fixed_steps = 0);
// whether the context's electrostatics can run as a fixed-length solve: conjugate
// gradients, rows checked after the evaluation, transforms by the library's own
// kernels, and no learned guess or guess history
bool elec_graph_ok(const mbpol_context* ctx);
// field evaluations for a fixed-length solve: the last solve's plus `extra`, at most max_iter + 1
int elec_graph_steps(const mbpol_context* ctx, int extra);
// whether the evaluation-level tuning may run trial evaluations (no learned guess or history)
bool elec_tune_ok(const mbpol_context* ctx);
// hold = true: keep the previous dipoles and solver state aside and start trial solves
// from alpha E; false: put them back, so the next evaluation runs as if no trial had
int elec_tune_hold(mbpol_context* ctx, bool hold);
// the host state a captured solve depends on (guess start, row capacities)
void elec_graph_key(const mbpol_context* ctx, long long key[4]);
// after a fixed-length solve has run: its statistics; `more` when it stopped before the
// stopping rule did (the evaluation is then run again with the host-driven solve)
int elec_graph_finish(mbpol_context* ctx, bool& more);
// after an evaluation: whether a neighbour row overflowed its capacity (the
// capacities have then grown and the evaluation must be run again)
int elec_rows_check(mbpol_context* ctx, bool& redo);

// what the learned dipole guess reads from the electrostatics of the last
// set-up (positions in the context's buffer)
struct ElecView {
    int n = 0, nw = 0;
    bool water_layout = false;  // 4 sites per water, O H H M, types 0 1 1 2
    const double *efix_pol = nullptr, *mu = nullptr;
    // history (kept while a learned guess of the previous-step kind exists): the converged
    // dipoles and alpha E_fixed of the last evaluation (prev) and of the one before it (prev2)
    const double *mu_prev = nullptr, *efix_prev = nullptr, *mu_prev2 = nullptr, *efix_prev2 = nullptr;
    bool have_previous = false, have_previous2 = false;
    const int* type = nullptr;
    const int *cell_of = nullptr, *cell_start = nullptr, *members = nullptr;
    int use_cells = 0, reach = 0;
    int3 nc{1, 1, 1};
    int periodic = 0;
    d3 box{0, 0, 0}, inv_box{0, 0, 0};
};
int elec_view(mbpol_context* ctx, ElecView& v);
// keep (on) or stop keeping the history of ElecView (counted: one per user)
void elec_keep_history(mbpol_context* ctx, bool on);
// learned guess (guess_driver.hip): mu0 for the polarisable sites from the
// context's attached model, after the electrostatics set-up of this evaluation
int guess_predict(mbpol_context* ctx, double* mu0);

// dispersion (disp_driver.hip)
int disp_create(mbpol_context* ctx);
void disp_destroy(mbpol_context* ctx);
int disp_set_box(mbpol_context* ctx);
int disp_evaluate(mbpol_context* ctx, hipStream_t st, u64* energy);
// per-group dispersion energies into out[g * 8 + 3]
int disp_group_sums(mbpol_context* ctx, hipStream_t st, u64* out);
// dispersion pairs within the cutoff at the last evaluation
int disp_pair_count(const mbpol_context* ctx, long long* pairs);

}  // namespace mbpol
