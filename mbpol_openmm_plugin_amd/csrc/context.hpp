// Host-side state of the MB-pol HIP platform: the system description (what
// the reference keeps in its Force objects) and the per-GPU context (what its
// Reference kernels keep between calls, MBPolReferenceKernels.h:45-244).
#pragma once
#include <hip/hip_runtime.h>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "kernels.h"

struct mbpol_system {
    int n = 0;
    double box[3] = {0, 0, 0};
    // ThreeParticleAverageSite
    std::vector<int> vs_idx;     // [site, p1, p2, p3]
    std::vector<double> vs_w;    // [w1, w2, w3]
    // MBPolOneBodyForce
    std::vector<int> ob_idx;
    int ob_method = 0;
    // MBPolTwoBodyForce
    std::vector<int> tb_idx;
    int tb_method = 0;
    double tb_cutoff = 1e10;
    // MBPolThreeBodyForce
    std::vector<int> th_idx;
    int th_method = 0;
    double th_cutoff = 1e10;
    // MBPolElectrostaticsForce
    int ne = 0;
    std::vector<double> el_q, el_damp, el_pol;
    std::vector<int> el_mol, el_type;
    int el_method = 0;
    double el_cutoff = 0.9, el_alpha = 0.0, el_tol = 1e-4, el_eps = 1e-7;
    int el_grid[3] = {0, 0, 0};
    int el_max_iter = 200;
    int el_redistribute = 1;
    double el_thole[5] = {0.4, 0.4, 0.055, 0.626, 0.055};
    // dispersion
    int nd = 0;
    std::vector<int> ds_type, ds_mol;
    int ds_ntypes = 0;
    std::vector<double> ds_c6, ds_d6;
    int ds_method = 0;
    double ds_cutoff = 1e10;
    // independent groups of waters (clusters), [grp_wstart[g], grp_wstart[g+1]); empty = one system
    std::vector<int> grp_wstart;
};

namespace mbpol {

struct Elec;
struct Disp;

// Grow-only device allocation.
// set while this thread captures an evaluation into a HIP graph: a buffer that
// would have to grow then fails instead (the capture is dropped and the
// evaluation runs eagerly)
extern thread_local bool g_capturing;

// A copy ordered with the context's work: enqueued on stream s and waited for.
// A plain hipMemcpy runs on the null stream, which the context's non-blocking
// streams do not wait for, and a host-to-device (pageable source) or
// device-to-device hipMemcpy may return before its data has landed.
inline hipError_t copy_on(hipStream_t s, void* dst, const void* src, size_t bytes) {
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s);
    return e != hipSuccess ? e : hipStreamSynchronize(s);
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf();
    hipError_t ensure(size_t nbytes);
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

// One O-O neighbour-list build (cell grid + pass buffers).
struct CellGrid {
    ListGeom g{};
    int ncell = 0;
    DevBuf cell_of, cell_count, cell_start, cell_fill, members;
};

// HIP-event timing of kernel groups on the context's stream.
// timed spans: whole phases, and single kernels of the electrostatics iteration
enum TimerCat { T_EVALUATE, T_ONEBODY, T_LISTS, T_TWOBODY, T_THREEBODY, T_DISPERSION, T_ELEC_SETUP, T_ELEC_ITER,
                T_ELEC_ENERGY, T_PME_SPREAD, T_FFT_Z_R2C, T_FFT_Y, T_FFT_X_CONV, T_FFT_Z_C2R, T_ELEC_ITERATE,
                T_FFT_PLANE_FWD, T_FFT_PLANE_INV, T_SPREAD_FFT_Z, T_NCAT };
struct Timer {
    bool enabled = false;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[T_NCAT];
    hipEvent_t open[T_NCAT] = {};
    double ms[T_NCAT] = {};
    long long count[T_NCAT] = {};
};

struct EvalStats {
    int iterations = 0;
    double epsilon = 0;
    int converged = 1;
    long long dimers = 0, trimers = 0, elec_pairs = 0, disp_pairs = 0, pol_pairs = 0;
};


// A host thread of the context's own that enqueues the forked short-range work
// while the calling thread enqueues the electrostatics, so that both streams are
// fed at once: a kernel launch costs the host several microseconds, and the
// electrostatics set-up and solve are many short kernels that the GPU runs
// faster than one thread enqueues them.  One job at a time: posted, then waited
// for before the join is enqueued.
struct SideEnqueuer {
    std::thread th;
    std::mutex m;
    std::condition_variable cv;
    std::function<int()> job;
    bool pending = false, quit = false;
    int rc = 0;
    std::string err;  // the job's error text (error texts are per thread)
};

}  // namespace mbpol

struct mbpol_guess;

struct mbpol_context {
    mbpol_system sys;
    int precision = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream for the short-range terms, forked from and joined to `stream` by events
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool evaluated = false;    // an evaluation has completed (launch shapes are tuned)
    int pip_blocks = 0;        // grid cap of the 2-body and 3-body kernels (0: one block per 64 units)
    int disp_blocks = 0;       // grid cap of the dispersion kernel (0: one wave per atom)
    int pip3_kernel = 0;       // 3-body kernel: 0 the polynomial through its relabelling group, 1 term by term
    int pip2_kernel = 1;       // 2-body kernel: the same choice (term by term by default)
    int pip3_blocks = 0;       // grid cap of the 3-body kernel (0: one block per 64 units)
    // the caps in waves per SIMD (blocks = waves x 4 x CUs), tuned on the first forked evaluation
    int pip_waves = 1, pip3_waves = 1, disp_waves = 0, cus = 0;
    // the short-range stream forks right after the virtual sites (1), after the set-up (0), or
    // right after the virtual sites with its work enqueued by the side enqueuer's thread (2)
    int fork_early = 2;
    // order of the 2-body, 3-body and dispersion work on the second stream, after its lists
    // (MBPOL_HIP_SR_ORDER, a permutation of "23d")
    char sr_order[4] = {'2', '3', 'd', 0};
    mbpol::SideEnqueuer* side_enq = nullptr;  // started on the first evaluation that uses it
    bool tuned = false;
    int tune_evaluations = 0;  // trial evaluations the tuning ran
    // the 1-body waters are the electrostatics waters (sites 4w..4w+3): the 1-body term
    // runs in the per-water kernel that computes the charges
    bool water_fused = false;
    bool onebody_with_charges = false;  // this evaluation's 1-body term runs with the charges
    // independent groups: water ranges, per-molecule candidate ranges, per-unit energies and per-group sums
    int ngroups = 0;
    mbpol::DevBuf g_wstart, g_mol_s, g_mol_e, g_eu1, g_eu2, g_eu3, g_energy;
    int full_chain_list = 0;
    mbpol::EvalStats stats;
    mbpol::Timer timer;

    // particle-level buffers
    mbpol::DevBuf pos, pos_resident, force_fx, force, energy_fx, vs_idx, vs_w;
    // pinned host staging
    void* h_counts = nullptr;
    // pinned staging of positions in and forces out for the host-buffer entry point
    double* h_pos = nullptr;
    double* h_force = nullptr;
    // scan scratch
    mbpol::DevBuf scan_tmp, scan_tmp_side;
    // 1B / 2B / 3B
    mbpol::DevBuf ob_idx, tb_idx, th_idx;
    mbpol::CellGrid grid2, grid3;
    mbpol::DevBuf dimer_count, dimer_off, dimers;
    long long dimer_cap = 0;
    mbpol::DevBuf lo_count, lo_start, lower;
    int lower_cap = 0;
    mbpol::DevBuf trip_count, trip_off, trips;
    long long trip_cap = 0;
    int lists_built = 0;        // lists built by this evaluation: 1 dimers, 2 trimers
    mbpol::DevBuf results_stage;  // energies and list lengths gathered for the one read-back (11 u64, every system)

    // HIP graph of the whole evaluation (mbpol_context_set_graph): captured on first
    // use, replayed while the host state it was captured from is unchanged
    // (graph_key; graph_gen counts the calls that change parameters)
    bool graph_on = false;
    hipGraphExec_t graph_exec = nullptr;
    long long graph_key[16] = {};
    long long graph_bad_key[16] = {};  // a key whose capture failed (evaluated eagerly)
    long long graph_gen = 0;
    int graph_steps = 0;               // field evaluations of the captured solve (0: from the last solve)
    int graph_lists_built = 0;
    bool fork_ready = false;           // a forked evaluation has run (every buffer of one is allocated)
    long long graph_captures = 0, graph_replays = 0, graph_fallbacks = 0;

    // electrostatics and dispersion live in their own structures
    mbpol::Elec* elec = nullptr;
    mbpol_guess* guess = nullptr;  // learned dipole guess in use (not owned)
    mbpol::Disp* disp = nullptr;
};
