// Host-visible declarations shared by the HIP kernel files and the context
// (host) code: small POD types and the launcher functions each kernel file
// exports.  No kernel is called across translation units directly.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mbpol {

// Start/stop events attached to one kernel launch (hipExtLaunchKernelGGL
// records them at the kernel's own start and end, so the span excludes the
// dispatch of the launch); both null when the launch is not timed.
struct KEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};

#define MBPOL_LAUNCH(ev, kernel, grid, block, shmem, stream, ...)                                            \
    do {                                                                                                     \
        if ((ev).start)                                                                                      \
            hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, (ev).start, (ev).stop, 0, __VA_ARGS__); \
        else                                                                                                 \
            hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);                             \
    } while (0)

struct d3 {
    double x, y, z;
};

typedef unsigned long long u64;

// geometry of one neighbour-list build
struct ListGeom {
    d3 L;          // box (nm); ignored when !periodic
    int periodic;
    int use_cells; // 27-cell search over an nc grid (periodic, nc >= 3), else all pairs
    int3 nc;
    double rc;     // nm
    // independent groups (non-periodic): candidates of molecule m are [gs[m], ge[m]); null = all molecules
    const int *gs, *ge;
};

// ---- kernels_water.hip ----
// per-group sums of unit energies into out[g * 8 + comp] (fixed point): group g covers the units
// [off[start[g]], off[start[g+1]]) when off is given, else [mult * start[g], mult * start[g+1])
void launch_group_sums(hipStream_t s, int ngroups, const int* start, int mult, const long long* off,
                       const double* eunit, u64* out, int comp, long long limit = 0x7fffffffffffffffLL);
// eunit (optional): energy of every unit (water, dimer, trimer) in kJ/mol, for per-group sums
void launch_place_vsites(hipStream_t s, int n, const int* vs_idx, const double* vs_w, double* pos);
// per-water terms: the 1-body energy/forces of waters ob[3w..3w+2] (into force/energy, eunit optional)
// and/or the charges of sites 4w..4w+3 with their derivatives (q, dq); with both, the waters coincide
void launch_water_terms(hipStream_t s, int nw, bool onebody, bool charges, const int* ob, const double* pos,
                        int periodic, d3 box, u64* force, u64* energy, double* eunit, double* q, double* dq);
void launch_forces_to_double(hipStream_t s, int n3, const u64* fx, double* f);
// zero `bytes` (a multiple of 4) with a kernel (no memset node in a captured graph)
void launch_fill_zero(hipStream_t s, void* p, size_t bytes);
// several buffers zeroed by one launch (lengths in bytes, multiples of 4): the
// buffers are gathered with add() and zeroed on stream s by launch(); a list
// that fills up is launched by add() and starts again
constexpr int kFillMax = 12;
struct FillBatch {
    int n = 0;
    void* p[kFillMax] = {};
    long long words[kFillMax] = {};
};
void launch_fill_multi(hipStream_t s, const FillBatch& B);
struct FillList {
    hipStream_t s;
    FillBatch B;
    explicit FillList(hipStream_t stream) : s(stream) {}
    void add(void* ptr, size_t bytes) {
        if (!ptr || bytes < 4) return;
        if (B.n == kFillMax) launch();
        B.p[B.n] = ptr;
        B.words[B.n++] = (long long)(bytes / 4);
    }
    void launch() {
        launch_fill_multi(s, B);
        B.n = 0;
    }
};
// energies (8) and the dimer, lower and trimer list lengths into out[0..10] (nullptr: not built)
void launch_collect_results(hipStream_t s, const unsigned long long* energy, const long long* nd, const int* nl,
                            const long long* nt, unsigned long long* out);
void launch_fold_vsites(hipStream_t s, int n, const int* vs_idx, const double* vs_w, double* f);

// ---- kernels_scan.hip: exclusive prefix sums (scratch of scan_scratch_bytes(n)) ----
size_t scan_scratch_bytes(long long n);
void launch_scan_i32(hipStream_t s, const int* in, int* out, long long n, void* scratch);
void launch_scan_i64(hipStream_t s, const long long* in, long long* out, long long n, void* scratch);

// ---- kernels_lists.hip ----
void launch_cell_count(hipStream_t s, int n, const int* idx, const double* pos, d3 L, int3 nc, int* cell_of,
                       int* cell_count);
void launch_cell_fill(hipStream_t s, int n, const int* cell_of, const int* cell_start, int* cell_fill,
                      int* cell_members);
void launch_dimer_pass(hipStream_t s, int n, const int* idx, const double* pos, ListGeom g, const int* cell_of,
                       const int* cell_start, const int* cell_members, long long* count, const long long* offset,
                       int2* pairs, long long cap);
void launch_lower_pass(hipStream_t s, int n, const int* idx, const double* pos, ListGeom g, const int* cell_of,
                       const int* cell_start, const int* cell_members, int* count, const int* offset, int* lower,
                       int cap);
void launch_trimer_pass(hipStream_t s, int n, const int* idx, const double* pos, ListGeom g, const int* lo_start,
                        const int* lower, int full_list, long long* count, const long long* offset, int* trip,
                        int lower_cap, long long cap);

// ---- kernels_pip.hip (one object per body order and precision) ----
// n: the list's capacity (grid size); ndev (optional): its length on the device (the kernels
// stop there); max_blocks > 0 caps the grid (64-lane blocks), the kernels then stride over the list
#define MBPOL_PIP_LAUNCHER(name, T)                                                                             \
    void name(hipStream_t s, long long n, const long long* ndev, const T* list, const int* idx, const double* pos, \
              int periodic, d3 box, u64* force, u64* energy, const KEvents& ev = KEvents(), double* eunit = nullptr, \
              int max_blocks = 0)
// 2-body, term by term (the coefficients as instruction literals)
MBPOL_PIP_LAUNCHER(launch_twobody_f32, int2);
MBPOL_PIP_LAUNCHER(launch_twobody_f64, int2);
// the 2-body polynomial through the dimer's relabelling group (pip_orbit_loop.h), FP32 and FP64
MBPOL_PIP_LAUNCHER(launch_twobody_orbit_f32, int2);
MBPOL_PIP_LAUNCHER(launch_twobody_orbit_f64, int2);
// FP64, term by term (the coefficients as instruction literals)
MBPOL_PIP_LAUNCHER(launch_threebody_f64, int);
// the polynomial through its relabelling group (pip_orbit_loop.h), FP32 and FP64
MBPOL_PIP_LAUNCHER(launch_threebody_orbit_f32, int);
MBPOL_PIP_LAUNCHER(launch_threebody_orbit_f64, int);
#undef MBPOL_PIP_LAUNCHER

}  // namespace mbpol
