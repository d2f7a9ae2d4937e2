// C ABI of the MB-pol HIP platform (include/mbpol_hip.h): system description,
// context set-up and the evaluation driver that sequences the kernels of one
// energy+force evaluation on the context's stream.
#include <algorithm>
#include <chrono>
#include <vector>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>

#include "../../include/mbpol_hip.h"
#include "context.hpp"
#include "elec.hpp"

using namespace mbpol;

// 3-body kernel (MBPOL_HIP_3B_KERNEL): "orbit" (default), the polynomial as 24
// passes of one orbit-representative polynomial over the trimer's relabelling
// group, coefficients read from a table (4096-water box: 626 us alone, 12.9x
// fewer L2 instruction requests); "terms", term by term with the coefficients
// as instruction literals (844 us).
static int pip3_kernel_choice() {
    const char* env = std::getenv("MBPOL_HIP_3B_KERNEL");
    return env && !std::strcmp(env, "terms") ? 1 : 0;
}
// the 2-body kernel: term by term unless MBPOL_HIP_2B_KERNEL=orbit (the loop over
// the dimer's relabelling group, 0.7% slower over a whole evaluation)
static int pip2_kernel_choice() {
    const char* env = std::getenv("MBPOL_HIP_2B_KERNEL");
    return env && !std::strcmp(env, "orbit") ? 0 : 1;
}

static thread_local std::string g_last_error;

int mbpol::set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t _e = (expr);                                                                         \
        if (_e != hipSuccess) return set_error(MBPOL_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

DevBuf::~DevBuf() {
    if (p) (void)hipFree(p);
}

thread_local bool mbpol::g_capturing = false;

hipError_t DevBuf::ensure(size_t nbytes) {
    if (nbytes <= bytes && p) return hipSuccess;
    if (mbpol::g_capturing) return hipErrorStreamCaptureUnsupported;
    if (p) {
        hipError_t e = hipFree(p);
        if (e != hipSuccess) return e;
        p = nullptr;
    }
    size_t want = std::max<size_t>(nbytes, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) bytes = want;
    else bytes = 0;
    return e;
}

// ---------------------------------------------------------------------------
// scans (kernels_scan.hip); each stream has its own scratch, so that scans on
// the two streams may overlap
// ---------------------------------------------------------------------------
static DevBuf& scan_scratch(mbpol_context* ctx, hipStream_t s) {
    return (s == ctx->stream) ? ctx->scan_tmp : ctx->scan_tmp_side;
}
int mbpol::scan_i32(mbpol_context* ctx, const int* in, int* out, int n, hipStream_t s) {
    if (!s) s = ctx->stream;
    DevBuf& tmp = scan_scratch(ctx, s);
    HIP_TRY(tmp.ensure(scan_scratch_bytes(n)));
    launch_scan_i32(s, in, out, n, tmp.p);
    return MBPOL_OK;
}
int mbpol::scan_i64(mbpol_context* ctx, const long long* in, long long* out, int n, hipStream_t s) {
    if (!s) s = ctx->stream;
    DevBuf& tmp = scan_scratch(ctx, s);
    HIP_TRY(tmp.ensure(scan_scratch_bytes(n)));
    launch_scan_i64(s, in, out, n, tmp.p);
    return MBPOL_OK;
}

// ---------------------------------------------------------------------------
// the side enqueuer: a host thread that runs one posted job at a time with the
// context's device current
// ---------------------------------------------------------------------------
static void side_enqueuer_loop(SideEnqueuer* w, int device) {
    const hipError_t dev = hipSetDevice(device);
    std::unique_lock<std::mutex> lk(w->m);
    for (;;) {
        w->cv.wait(lk, [w] { return w->pending || w->quit; });
        if (!w->pending) return;
        std::function<int()> f = std::move(w->job);
        lk.unlock();
        const int e = dev == hipSuccess ? f() : set_error(MBPOL_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(dev));
        std::string msg = e ? g_last_error : std::string();
        lk.lock();
        w->rc = e;
        w->err = std::move(msg);
        w->pending = false;
        w->cv.notify_all();
    }
}

static int side_post(mbpol_context* ctx, std::function<int()> f) {
    if (!ctx->side_enq) {
        auto* w = new SideEnqueuer;
        try {
            w->th = std::thread(side_enqueuer_loop, w, ctx->device);
        } catch (const std::exception& ex) {
            delete w;
            return set_error(MBPOL_ERR_HIP, "side enqueuer thread: %s", ex.what());
        }
        ctx->side_enq = w;
    }
    SideEnqueuer* w = ctx->side_enq;
    std::lock_guard<std::mutex> lk(w->m);
    w->job = std::move(f);
    w->pending = true;
    w->cv.notify_all();
    return MBPOL_OK;
}

// waits for the posted job; its error (code and text) becomes this thread's
static int side_wait(mbpol_context* ctx) {
    SideEnqueuer* w = ctx->side_enq;
    std::unique_lock<std::mutex> lk(w->m);
    w->cv.wait(lk, [w] { return !w->pending; });
    return w->rc ? set_error(w->rc, "%s", w->err.c_str()) : MBPOL_OK;
}

static void side_stop(mbpol_context* ctx) {
    SideEnqueuer* w = ctx->side_enq;
    if (!w) return;
    {
        std::lock_guard<std::mutex> lk(w->m);
        w->quit = true;
        w->cv.notify_all();
    }
    w->th.join();
    delete w;
    ctx->side_enq = nullptr;
}

// ---------------------------------------------------------------------------
// HIP-event timing
// ---------------------------------------------------------------------------
static hipEvent_t take_event(mbpol_context* ctx) {
    Timer& t = ctx->timer;
    if (!t.pool.empty()) {
        hipEvent_t e = t.pool.back();
        t.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void mbpol::timer_start(mbpol_context* ctx, int cat) {
    if (!ctx->timer.enabled) return;
    hipEvent_t e = take_event(ctx);
    if (!e) return;
    (void)hipEventRecord(e, ctx->stream);
    ctx->timer.open[cat] = e;
}

mbpol::KEvents mbpol::timer_kernel(mbpol_context* ctx, int cat) {
    (void)cat;
    KEvents ev;
    if (!ctx->timer.enabled) return ev;
    ev.start = take_event(ctx);
    ev.stop = ev.start ? take_event(ctx) : nullptr;
    if (!ev.stop) ev.start = nullptr;
    return ev;
}
void mbpol::timer_kernel_done(mbpol_context* ctx, int cat, const KEvents& ev) {
    if (ev.start && ev.stop) ctx->timer.pending[cat].push_back({ev.start, ev.stop});
}
void mbpol::timer_stop(mbpol_context* ctx, int cat) {
    Timer& t = ctx->timer;
    if (!t.enabled || !t.open[cat]) return;
    hipEvent_t e = take_event(ctx);
    if (!e) return;
    (void)hipEventRecord(e, ctx->stream);
    t.pending[cat].push_back({t.open[cat], e});
    t.open[cat] = nullptr;
}

static void timer_collect(mbpol_context* ctx) {
    Timer& t = ctx->timer;
    (void)hipStreamSynchronize(ctx->stream);
    for (int c = 0; c < T_NCAT; ++c) {
        for (auto& pr : t.pending[c]) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
                t.ms[c] += ms;
                t.count[c] += 1;
            }
            t.pool.push_back(pr.first);
            t.pool.push_back(pr.second);
        }
        t.pending[c].clear();
    }
}

// ---------------------------------------------------------------------------
// system description
// ---------------------------------------------------------------------------
extern "C" {

const char* mbpol_last_error(void) { return g_last_error.c_str(); }

int mbpol_device_count(int32_t* count) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    *count = c;
    return MBPOL_OK;
}

int mbpol_system_create(int32_t num_particles, mbpol_system** out) {
    if (!out || num_particles <= 0) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "num_particles must be positive");
    mbpol_system* s = new mbpol_system();
    s->n = num_particles;
    *out = s;
    return MBPOL_OK;
}

void mbpol_system_destroy(mbpol_system* sys) { delete sys; }

int mbpol_system_set_box(mbpol_system* sys, const double box_nm[3]) {
    for (int d = 0; d < 3; ++d) {
        if (box_nm[d] < 0) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "negative box length");
        sys->box[d] = box_nm[d];
    }
    return MBPOL_OK;
}

int mbpol_system_set_groups(mbpol_system* sys, int32_t ngroups, const int32_t* water_start) {
    if (!sys || ngroups < 0 || (ngroups > 0 && !water_start)) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "bad argument");
    sys->grp_wstart.clear();
    if (ngroups == 0) return MBPOL_OK;
    if (water_start[0] != 0) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "groups: the first group starts at water 0");
    for (int g = 0; g < ngroups; ++g)
        if (water_start[g + 1] <= water_start[g])
            return set_error(MBPOL_ERR_INVALID_ARGUMENT, "groups: group %d is empty or out of order", g);
    sys->grp_wstart.assign(water_start, water_start + ngroups + 1);
    return MBPOL_OK;
}

static bool valid_particle(const mbpol_system* s, int i) { return i >= 0 && i < s->n; }

int mbpol_system_add_virtual_site(mbpol_system* sys, int32_t site, int32_t p1, int32_t p2, int32_t p3, double w1,
                                  double w2, double w3) {
    if (!valid_particle(sys, site) || !valid_particle(sys, p1) || !valid_particle(sys, p2) || !valid_particle(sys, p3))
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "virtual site index out of range");
    sys->vs_idx.insert(sys->vs_idx.end(), {site, p1, p2, p3});
    sys->vs_w.insert(sys->vs_w.end(), {w1, w2, w3});
    return MBPOL_OK;
}

static int check_molecules(const mbpol_system* sys, int32_t n, const int32_t* idx, const char* what) {
    if (n < 0 || (n > 0 && !idx)) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "%s: bad molecule list", what);
    for (int i = 0; i < 3 * n; ++i)
        if (!valid_particle(sys, idx[i]))
            return set_error(MBPOL_ERR_INVALID_ARGUMENT, "%s: particle index %d out of range", what, idx[i]);
    return MBPOL_OK;
}

int mbpol_system_set_onebody(mbpol_system* sys, int32_t n, const int32_t* idx, int32_t method) {
    if (int rc = check_molecules(sys, n, idx, "MBPolOneBodyForce")) return rc;
    if (method != MBPOL_ONEBODY_NONPERIODIC && method != MBPOL_ONEBODY_PERIODIC)
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "MBPolOneBodyForce: unknown nonbonded method %d", method);
    sys->ob_idx.assign(idx, idx + 3 * n);
    sys->ob_method = method;
    return MBPOL_OK;
}

static int set_nbody(mbpol_system* sys, int32_t n, const int32_t* idx, int32_t method, double cutoff,
                     std::vector<int>& dst, int& m, double& rc, const char* what) {
    if (int e = check_molecules(sys, n, idx, what)) return e;
    if (method < 0 || method > 2) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "%s: unknown nonbonded method %d", what, method);
    if (!(cutoff > 0)) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "%s: cutoff must be positive", what);
    dst.assign(idx, idx + 3 * n);
    m = method;
    rc = cutoff;
    return MBPOL_OK;
}

int mbpol_system_set_twobody(mbpol_system* sys, int32_t n, const int32_t* idx, int32_t method, double cutoff_nm) {
    return set_nbody(sys, n, idx, method, cutoff_nm, sys->tb_idx, sys->tb_method, sys->tb_cutoff, "MBPolTwoBodyForce");
}

int mbpol_system_set_threebody(mbpol_system* sys, int32_t n, const int32_t* idx, int32_t method, double cutoff_nm) {
    return set_nbody(sys, n, idx, method, cutoff_nm, sys->th_idx, sys->th_method, sys->th_cutoff,
                     "MBPolThreeBodyForce");
}

int mbpol_system_set_electrostatics(mbpol_system* sys, int32_t n, const double* charge, const int32_t* molecule,
                                    const int32_t* atom_type, const double* damping, const double* polarity,
                                    int32_t method, double cutoff_nm, double ewald_alpha, const int32_t grid[3],
                                    double ewald_tol, int32_t max_iter, double target_eps,
                                    int32_t charge_redistribution, const double thole[5]) {
    // MBPolElectrostaticsForceImpl.cpp:53-54: one entry per particle of the System
    if (n != sys->n)
        return set_error(MBPOL_ERR_INVALID_ARGUMENT,
                         "MBPolElectrostaticsForce must have exactly as many particles as the System it belongs to.");
    if (method != MBPOL_ELEC_NOCUTOFF && method != MBPOL_ELEC_PME)
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "MBPolElectrostaticsForce: unknown nonbonded method %d", method);
    if (charge_redistribution && n % 4)
        return set_error(MBPOL_ERR_INVALID_ARGUMENT,
                         "charge redistribution needs O,H1,H2,M sites per water (4 per molecule)");
    for (int i = 0; i < n; ++i)
        if (atom_type[i] < 0 || atom_type[i] > 2)
            return set_error(MBPOL_ERR_INVALID_ARGUMENT, "atom type must be 0 (O), 1 (H) or 2 (M)");
    sys->ne = n;
    sys->el_q.assign(charge, charge + n);
    sys->el_mol.assign(molecule, molecule + n);
    sys->el_type.assign(atom_type, atom_type + n);
    sys->el_damp.assign(damping, damping + n);
    sys->el_pol.assign(polarity, polarity + n);
    sys->el_method = method;
    sys->el_cutoff = cutoff_nm;
    sys->el_alpha = ewald_alpha;
    for (int d = 0; d < 3; ++d) sys->el_grid[d] = grid ? grid[d] : 0;
    sys->el_tol = ewald_tol;
    sys->el_max_iter = max_iter;
    sys->el_eps = target_eps;
    sys->el_redistribute = charge_redistribution;
    if (thole)
        for (int k = 0; k < 5; ++k) sys->el_thole[k] = thole[k];
    return MBPOL_OK;
}

int mbpol_system_set_dispersion(mbpol_system* sys, int32_t n, const int32_t* type, const int32_t* molecule,
                                int32_t num_types, const double* c6, const double* d6, int32_t method,
                                double cutoff_nm) {
    if (n != sys->n) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "dispersion: one type per particle required");
    if (num_types <= 0 || num_types > 16 || method < 0 || method > 2)
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "dispersion: bad type count (1..16) or method");
    for (int i = 0; i < n; ++i)
        if (molecule[i] < 0 || molecule[i] >= (1 << 26))
            return set_error(MBPOL_ERR_INVALID_ARGUMENT, "dispersion: molecule index out of range");
    for (int i = 0; i < n; ++i)
        if (type[i] < 0 || type[i] >= num_types) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "dispersion: bad type");
    sys->nd = n;
    sys->ds_type.assign(type, type + n);
    sys->ds_mol.assign(molecule, molecule + n);
    sys->ds_ntypes = num_types;
    sys->ds_c6.assign(c6, c6 + num_types * num_types);
    sys->ds_d6.assign(d6, d6 + num_types * num_types);
    sys->ds_method = method;
    sys->ds_cutoff = cutoff_nm;
    return MBPOL_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
static bool periodic_box(const mbpol_system& s) { return s.box[0] > 0 && s.box[1] > 0 && s.box[2] > 0; }

template <typename T>
static int upload(mbpol_context* ctx, DevBuf& b, const std::vector<T>& v) {
    HIP_TRY(b.ensure(std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty()) HIP_TRY(copy_on(ctx->stream, b.p, v.data(), v.size() * sizeof(T)));
    return MBPOL_OK;
}

// box/cutoff checks of the reference (MBPol*ForceImpl.cpp:60-66, MBPolReferenceKernels.cpp:224-227,387-390,471-474)
static int check_box(const mbpol_system& s) {
    const bool pbc = periodic_box(s);
    auto chk = [&](bool uses_pbc, double rc, const char* what) -> int {
        if (!uses_pbc) return MBPOL_OK;
        if (!pbc) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "%s: periodic method without a periodic box", what);
        for (int d = 0; d < 3; ++d)
            if (s.box[d] < 1.999999 * rc)
                return set_error(MBPOL_ERR_BOX_TOO_SMALL,
                                 "%s: The periodic box size has decreased to less than twice the cutoff.", what);
        return MBPOL_OK;
    };
    if (int e = chk(!s.tb_idx.empty() && s.tb_method == MBPOL_CUTOFF_PERIODIC, s.tb_cutoff, "MBPolTwoBodyForce")) return e;
    if (int e = chk(!s.th_idx.empty() && s.th_method == MBPOL_CUTOFF_PERIODIC, s.th_cutoff, "MBPolThreeBodyForce")) return e;
    if (int e = chk(s.ne > 0 && s.el_method == MBPOL_ELEC_PME, s.el_cutoff, "MBPolElectrostaticsForce")) return e;
    if (int e = chk(s.nd > 0 && s.ds_method == MBPOL_CUTOFF_PERIODIC, s.ds_cutoff, "dispersion")) return e;
    if (!s.ob_idx.empty() && s.ob_method == MBPOL_ONEBODY_PERIODIC && !pbc)
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "MBPolOneBodyForce: periodic method without a periodic box");
    return MBPOL_OK;
}

// Cell grid for an O-O list: 27-cell search when the periodic box holds >= 3 cells of edge >= rc per
// dimension, otherwise every pair is visited (small clusters, huge cutoffs).
static int setup_grid(mbpol_context* ctx, CellGrid& cg, int n, bool periodic, double rc) {
    const mbpol_system& s = ctx->sys;
    cg.g.periodic = periodic ? 1 : 0;
    cg.g.L = d3{s.box[0], s.box[1], s.box[2]};
    cg.g.rc = rc;
    cg.g.use_cells = 0;
    cg.g.nc = make_int3(1, 1, 1);
    if (periodic) {
        int nc[3];
        for (int d = 0; d < 3; ++d) nc[d] = (int)std::floor(s.box[d] / rc);
        if (nc[0] >= 3 && nc[1] >= 3 && nc[2] >= 3) {
            cg.g.use_cells = 1;
            cg.g.nc = make_int3(nc[0], nc[1], nc[2]);
        }
    }
    cg.ncell = cg.g.nc.x * cg.g.nc.y * cg.g.nc.z;
    HIP_TRY(cg.cell_of.ensure(sizeof(int) * std::max(n, 1)));
    HIP_TRY(cg.cell_count.ensure(sizeof(int) * (cg.ncell + 1)));
    HIP_TRY(cg.cell_start.ensure(sizeof(int) * (cg.ncell + 1)));
    HIP_TRY(cg.cell_fill.ensure(sizeof(int) * (cg.ncell + 1)));
    HIP_TRY(cg.members.ensure(sizeof(int) * std::max(n, 1)));
    return MBPOL_OK;
}

// Independent groups (clusters) of waters in one non-periodic system: every
// water occupies particles 4w..4w+3 (O, H1, H2, M), the 1-, 2- and 3-body
// molecule lists are the waters in order, and no Force uses a periodic method.
static int check_groups(const mbpol_system& s) {
    const int W = s.grp_wstart.back();
    auto bad = [&](const char* what) { return set_error(MBPOL_ERR_INVALID_ARGUMENT, "groups: %s", what); };
    if (s.n != 4 * W) return bad("the system must hold exactly 4 particles (O, H1, H2, M) per water of the groups");
    if (periodic_box(s)) return bad("groups need a non-periodic system");
    auto waters = [&](const std::vector<int>& idx) {
        if (idx.empty()) return true;
        if ((int)idx.size() != 3 * W) return false;
        for (int w = 0; w < W; ++w)
            if (idx[3 * w] != 4 * w || idx[3 * w + 1] != 4 * w + 1 || idx[3 * w + 2] != 4 * w + 2) return false;
        return true;
    };
    if (!waters(s.ob_idx) || !waters(s.tb_idx) || !waters(s.th_idx))
        return bad("the 1-, 2- and 3-body molecule lists must be the waters in order");
    if ((!s.tb_idx.empty() && s.tb_method == MBPOL_CUTOFF_PERIODIC) ||
        (!s.th_idx.empty() && s.th_method == MBPOL_CUTOFF_PERIODIC) ||
        (s.nd > 0 && s.ds_method == MBPOL_CUTOFF_PERIODIC) || (s.ne > 0 && s.el_method == MBPOL_ELEC_PME))
        return bad("periodic methods cannot be combined with groups");
    if (s.ne > 0 && (s.ne != 4 * W || !s.el_redistribute)) return bad("electrostatics must cover the O H H M sites of every water");
    if (s.nd > 0)
        for (int i = 0; i < s.nd; ++i)
            if (s.ds_mol[i] != i / 4) return bad("dispersion molecules must be the waters");
    return MBPOL_OK;
}

// device arrays of the groups and the candidate ranges of the molecule lists
static int groups_create(mbpol_context* ctx) {
    const std::vector<int>& ws = ctx->sys.grp_wstart;
    ctx->ngroups = ws.empty() ? 0 : (int)ws.size() - 1;
    if (!ctx->ngroups) return MBPOL_OK;
    const int W = ws.back();
    std::vector<int> ms(W), me(W);
    for (int g = 0; g < ctx->ngroups; ++g)
        for (int w = ws[g]; w < ws[g + 1]; ++w) {
            ms[w] = ws[g];
            me[w] = ws[g + 1];
        }
    if (int e = upload(ctx, ctx->g_wstart, ws)) return e;
    if (int e = upload(ctx, ctx->g_mol_s, ms)) return e;
    if (int e = upload(ctx, ctx->g_mol_e, me)) return e;
    HIP_TRY(ctx->g_eu1.ensure(sizeof(double) * std::max(W, 1)));
    HIP_TRY(ctx->g_energy.ensure(sizeof(u64) * 8 * ctx->ngroups));
    return MBPOL_OK;
}

static void apply_list_groups(mbpol_context* ctx) {
    for (CellGrid* cg : {&ctx->grid2, &ctx->grid3}) {
        cg->g.gs = ctx->ngroups ? ctx->g_mol_s.as<int>() : nullptr;
        cg->g.ge = ctx->ngroups ? ctx->g_mol_e.as<int>() : nullptr;
    }
}

static int bin_cells(mbpol_context* ctx, hipStream_t st, CellGrid& cg, int n, const int* idx) {
    if (!cg.g.use_cells) return MBPOL_OK;
    launch_fill_zero(st, cg.cell_count.p, sizeof(int) * (cg.ncell + 1));
    launch_fill_zero(st, cg.cell_fill.p, sizeof(int) * (cg.ncell + 1));
    launch_cell_count(st, n, idx, ctx->pos.as<double>(), cg.g.L, cg.g.nc, cg.cell_of.as<int>(),
                      cg.cell_count.as<int>());
    if (int e = scan_i32(ctx, cg.cell_count.as<int>(), cg.cell_start.as<int>(), cg.ncell + 1, st)) return e;
    launch_cell_fill(st, n, cg.cell_of.as<int>(), cg.cell_start.as<int>(), cg.cell_fill.as<int>(),
                     cg.members.as<int>());
    return MBPOL_OK;
}

extern "C" {

int mbpol_context_create(const mbpol_system* sys, int32_t precision, int32_t device, mbpol_context** out) {
    if (sys && !sys->grp_wstart.empty())
        if (int e = check_groups(*sys)) return e;
    if (!sys || !out) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null argument");
    if (precision != MBPOL_PRECISION_MIXED && precision != MBPOL_PRECISION_DOUBLE &&
        precision != MBPOL_PRECISION_SINGLE_PIP)
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "unknown precision mode %d", precision);
    int ndev = 0;
    const hipError_t derr = hipGetDeviceCount(&ndev);
    if (derr != hipSuccess || ndev <= 0)
        return set_error(MBPOL_ERR_NO_DEVICE, "no HIP device available (%s)",
                         derr != hipSuccess ? hipGetErrorString(derr) : "0 devices");
    if (device < 0 || device >= ndev) return set_error(MBPOL_ERR_NO_DEVICE, "device %d not present", device);
    // cutoff <= half box at initialisation (MBPolTwoBodyForceImpl.cpp:60-66 and siblings)
    if (periodic_box(*sys)) {
        auto half = [&](bool on, double rc, const char* what) -> int {
            if (!on) return MBPOL_OK;
            for (int d = 0; d < 3; ++d)
                if (rc > 0.5 * sys->box[d])
                    return set_error(MBPOL_ERR_CUTOFF_TOO_LARGE,
                                     "%s: The cutoff distance cannot be greater than half the periodic box size.", what);
            return MBPOL_OK;
        };
        if (int e = half(!sys->tb_idx.empty() && sys->tb_method == MBPOL_CUTOFF_PERIODIC, sys->tb_cutoff, "MBPolTwoBodyForce")) return e;
        if (int e = half(!sys->th_idx.empty() && sys->th_method == MBPOL_CUTOFF_PERIODIC, sys->th_cutoff, "MBPolThreeBodyForce")) return e;
        if (int e = half(sys->ne > 0 && sys->el_method == MBPOL_ELEC_PME, sys->el_cutoff, "MBPolElectrostaticsForce")) return e;
    }
    if (int e = check_box(*sys)) return e;
    HIP_TRY(hipSetDevice(device));
    mbpol_context* ctx = new mbpol_context();
    ctx->sys = *sys;
    ctx->precision = precision;
    ctx->device = device;
    auto fail = [&](int code) {
        mbpol_context_destroy(ctx);
        return code;
    };
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess)
        return fail(set_error(MBPOL_ERR_HIP, "hipStreamCreate failed"));
    if (hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess)
        return fail(set_error(MBPOL_ERR_HIP, "hipEventCreate failed"));
    if (hipHostMalloc(&ctx->h_counts, 4096, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_pos, sizeof(double) * 3 * std::max(sys->n, 1), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_force, sizeof(double) * 3 * std::max(sys->n, 1), hipHostMallocDefault) != hipSuccess)
        return fail(set_error(MBPOL_ERR_HIP, "hipHostMalloc failed"));
    {
        // Grid cap of the 2-body and 3-body kernels: one wave per SIMD of the device
        // (MBPOL_HIP_PIP_WAVES: waves per SIMD, 0 = one block per 64 units).  The
        // whole capped grid is resident at once, so the dispatcher holds no queue of
        // their blocks and the electrostatics kernels on the other stream are
        // dispatched as soon as they are launched.
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess)
            return fail(set_error(MBPOL_ERR_HIP, "hipGetDeviceProperties failed"));
        int waves = 1;
        if (const char* env = std::getenv("MBPOL_HIP_PIP_WAVES")) waves = std::max(0, std::atoi(env));
        ctx->cus = prop.multiProcessorCount;
        ctx->pip_waves = waves;
        ctx->pip_blocks = waves * 4 * prop.multiProcessorCount;
        // the 3-body kernel and its grid cap (MBPOL_HIP_3B_WAVES waves per SIMD): one wave
        // per SIMD, so that its 256-VGPR waves leave every SIMD room for the electrostatics
        // kernels (uncapped, they hold every SIMD's registers and the next electrostatics
        // launch waits for them to retire); tuned with the other caps
        ctx->pip3_kernel = pip3_kernel_choice();
        ctx->pip2_kernel = pip2_kernel_choice();
        int waves3 = 1;
        if (const char* env = std::getenv("MBPOL_HIP_3B_WAVES")) waves3 = std::max(0, std::atoi(env));
        ctx->pip3_waves = waves3;
        ctx->pip3_blocks = waves3 * 4 * prop.multiProcessorCount;
        int dwaves = 0;  // the dispersion kernel: one block per 64 atoms (MBPOL_HIP_DISP_WAVES)
        if (const char* env = std::getenv("MBPOL_HIP_DISP_WAVES")) dwaves = std::max(0, std::atoi(env));
        ctx->disp_waves = dwaves;
        if (const char* env = std::getenv("MBPOL_HIP_SR_ORDER")) {  // a permutation of "23d"
            std::string o(env), sorted_o(env);
            std::sort(sorted_o.begin(), sorted_o.end());
            if (sorted_o == "23d") std::memcpy(ctx->sr_order, o.c_str(), 4);
        }
        if (const char* env = std::getenv("MBPOL_HIP_FORK"))
            ctx->fork_early = !std::strcmp(env, "late") ? 0 : !std::strcmp(env, "thread") ? 2 : 1;
        ctx->disp_blocks = dwaves * 4 * prop.multiProcessorCount;
    }
    const int n = sys->n;
    if (ctx->pos.ensure(sizeof(double) * 3 * n) || ctx->force_fx.ensure(sizeof(u64) * 3 * n) ||
        ctx->force.ensure(sizeof(double) * 3 * n) || ctx->energy_fx.ensure(sizeof(u64) * 8) ||
        ctx->results_stage.ensure(sizeof(u64) * 16))
        return fail(set_error(MBPOL_ERR_HIP, "device allocation failed"));
    if (int e = upload(ctx, ctx->vs_idx, sys->vs_idx)) return fail(e);
    if (int e = upload(ctx, ctx->vs_w, sys->vs_w)) return fail(e);
    if (int e = upload(ctx, ctx->ob_idx, sys->ob_idx)) return fail(e);
    if (int e = upload(ctx, ctx->tb_idx, sys->tb_idx)) return fail(e);
    if (int e = upload(ctx, ctx->th_idx, sys->th_idx)) return fail(e);
    if (int e = groups_create(ctx)) return fail(e);
    const int n2 = (int)sys->tb_idx.size() / 3, n3 = (int)sys->th_idx.size() / 3;
    if (n2 > 0) {
        if (int e = setup_grid(ctx, ctx->grid2, n2, sys->tb_method == MBPOL_CUTOFF_PERIODIC, sys->tb_cutoff)) return fail(e);
        if (ctx->dimer_count.ensure(sizeof(long long) * (n2 + 1)) || ctx->dimer_off.ensure(sizeof(long long) * (n2 + 1)))
            return fail(set_error(MBPOL_ERR_HIP, "device allocation failed"));
    }
    if (n3 > 0) {
        if (int e = setup_grid(ctx, ctx->grid3, n3, sys->th_method == MBPOL_CUTOFF_PERIODIC, sys->th_cutoff)) return fail(e);
        if (ctx->lo_count.ensure(sizeof(int) * (n3 + 1)) || ctx->lo_start.ensure(sizeof(int) * (n3 + 1)) ||
            ctx->trip_count.ensure(sizeof(long long) * (n3 + 1)) || ctx->trip_off.ensure(sizeof(long long) * (n3 + 1)))
            return fail(set_error(MBPOL_ERR_HIP, "device allocation failed"));
    }
    apply_list_groups(ctx);
    if (sys->ne > 0) {
        if (int e = elec_create(ctx)) return fail(e);
        // the per-water kernel computes the 1-body term with the charges when the
        // 1-body waters are the electrostatics waters, in order
        const int nw1 = (int)sys->ob_idx.size() / 3;
        bool same = sys->el_redistribute && nw1 > 0 && 4 * nw1 == sys->ne;
        for (int w = 0; same && w < nw1; ++w)
            same = sys->ob_idx[3 * w] == 4 * w && sys->ob_idx[3 * w + 1] == 4 * w + 1 && sys->ob_idx[3 * w + 2] == 4 * w + 2;
        ctx->water_fused = same;
    }
    if (sys->nd > 0) {
        if (int e = disp_create(ctx)) return fail(e);
    }
    *out = ctx;
    return MBPOL_OK;
}

void mbpol_context_destroy(mbpol_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    side_stop(ctx);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    timer_collect(ctx);
    for (hipEvent_t e : ctx->timer.pool) (void)hipEventDestroy(e);
    elec_destroy(ctx);
    disp_destroy(ctx);
    if (ctx->h_counts) (void)hipHostFree(ctx->h_counts);
    if (ctx->h_pos) (void)hipHostFree(ctx->h_pos);
    if (ctx->h_force) (void)hipHostFree(ctx->h_force);
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int mbpol_context_set_box(mbpol_context* ctx, const double box_nm[3]) {
    if (ctx) ++ctx->graph_gen;
    mbpol_system s = ctx->sys;
    for (int d = 0; d < 3; ++d) s.box[d] = box_nm[d];
    if (int e = check_box(s)) return e;
    for (int d = 0; d < 3; ++d) ctx->sys.box[d] = box_nm[d];
    HIP_TRY(hipSetDevice(ctx->device));
    const int n2 = (int)ctx->sys.tb_idx.size() / 3, n3 = (int)ctx->sys.th_idx.size() / 3;
    if (n2 > 0)
        if (int e = setup_grid(ctx, ctx->grid2, n2, ctx->sys.tb_method == MBPOL_CUTOFF_PERIODIC, ctx->sys.tb_cutoff)) return e;
    if (n3 > 0)
        if (int e = setup_grid(ctx, ctx->grid3, n3, ctx->sys.th_method == MBPOL_CUTOFF_PERIODIC, ctx->sys.th_cutoff)) return e;
    apply_list_groups(ctx);
    if (ctx->elec)
        if (int e = elec_set_box(ctx)) return e;
    if (ctx->disp)
        if (int e = disp_set_box(ctx)) return e;
    return MBPOL_OK;
}

// copyParametersToContext of the 1-, 2- and 3-body Forces
// (MBPolReferenceKernels.cpp:116-127, 401-414, 485-498): same molecule count, new index lists
static int update_molecules(mbpol_context* ctx, int32_t n, const int32_t* idx, std::vector<int>& host, DevBuf& dev,
                            const char* what, const char* count_msg) {
    if (ctx) ++ctx->graph_gen;
    if (!ctx) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null context");
    if (n != (int32_t)(host.size() / 3)) return set_error(MBPOL_ERR_COUNT_CHANGED, "updateParametersInContext: %s", count_msg);
    if (int e = check_molecules(&ctx->sys, n, idx, what)) return e;
    const std::vector<int> old = host;
    host.assign(idx, idx + 3 * n);
    if (!ctx->sys.grp_wstart.empty())
        if (int e = check_groups(ctx->sys)) {
            host = old;
            return e;
        }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->side) HIP_TRY(hipStreamSynchronize(ctx->side));
    if (n > 0) HIP_TRY(copy_on(ctx->stream, dev.p, host.data(), sizeof(int) * host.size()));
    return MBPOL_OK;
}

int mbpol_context_update_onebody(mbpol_context* ctx, int32_t n, const int32_t* idx) {
    if (!ctx) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null context");
    if (int e = update_molecules(ctx, n, idx, ctx->sys.ob_idx, ctx->ob_idx, "MBPolOneBodyForce",
                                 "The number of stretch-bends has changed"))
        return e;
    const mbpol_system& s = ctx->sys;
    bool same = ctx->elec && s.el_redistribute && 4 * n == s.ne;
    for (int w = 0; same && w < n; ++w)
        same = s.ob_idx[3 * w] == 4 * w && s.ob_idx[3 * w + 1] == 4 * w + 1 && s.ob_idx[3 * w + 2] == 4 * w + 2;
    ctx->water_fused = same;
    return MBPOL_OK;
}

int mbpol_context_update_twobody(mbpol_context* ctx, int32_t n, const int32_t* idx) {
    if (!ctx) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null context");
    return update_molecules(ctx, n, idx, ctx->sys.tb_idx, ctx->tb_idx, "MBPolTwoBodyForce",
                            "The number of particles has changed");
}

int mbpol_context_update_threebody(mbpol_context* ctx, int32_t n, const int32_t* idx) {
    if (!ctx) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null context");
    return update_molecules(ctx, n, idx, ctx->sys.th_idx, ctx->th_idx, "MBPolThreeBodyForce",
                            "The number of particles has changed");
}

int mbpol_context_set_list_mode(mbpol_context* ctx, int32_t full_chain_list) {
    ++ctx->graph_gen;
    ctx->full_chain_list = full_chain_list ? 1 : 0;
    return MBPOL_OK;
}

int mbpol_context_set_graph(mbpol_context* ctx, int32_t enabled) {
    if (!ctx) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null context");
    ctx->graph_on = enabled != 0;
    if (!ctx->graph_on && ctx->graph_exec) {
        HIP_TRY(hipSetDevice(ctx->device));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(hipGraphExecDestroy(ctx->graph_exec));
        ctx->graph_exec = nullptr;
    }
    return MBPOL_OK;
}

}  // extern "C"

// Build the dimer list and the trimer work list.  The list lengths stay on the
// device (dimer_off[n2], lo_start[n3], trip_off[n3]): the fill passes write
// what fits in the current capacities, the PIP kernels read the lengths, and
// the host compares lengths and capacities after the evaluation (list_check),
// growing the lists and running the evaluation again when one did not fit.
// Before the first evaluation (no capacities yet) the lengths are read back in
// between the passes, which sizes the lists.
static int build_lists(mbpol_context* ctx, hipStream_t st, bool want2, bool want3) {
    const mbpol_system& s = ctx->sys;
    const int n2 = (int)s.tb_idx.size() / 3, n3 = (int)s.th_idx.size() / 3;
    const double* pos = ctx->pos.as<double>();
    long long* hc = static_cast<long long*>(ctx->h_counts);
    auto grow2 = [&](long long nd) -> int {
        if (nd > ctx->dimer_cap) {
            HIP_TRY(ctx->dimers.ensure(sizeof(int2) * (nd + nd / 4 + 64)));
            ctx->dimer_cap = nd + nd / 4 + 64;
        }
        return MBPOL_OK;
    };
    if (want2) {
        CellGrid& g = ctx->grid2;
        if (int e = bin_cells(ctx, st, g, n2, ctx->tb_idx.as<int>())) return e;
        launch_fill_zero(st, ctx->dimer_count.p, sizeof(long long) * (n2 + 1));
        launch_dimer_pass(st, n2, ctx->tb_idx.as<int>(), pos, g.g, g.cell_of.as<int>(), g.cell_start.as<int>(),
                          g.members.as<int>(), ctx->dimer_count.as<long long>(), nullptr, nullptr, 0);
        if (int e = scan_i64(ctx, ctx->dimer_count.as<long long>(), ctx->dimer_off.as<long long>(), n2 + 1, st)) return e;
        if (ctx->dimer_cap == 0) {  // first build: size the list
            HIP_TRY(hipMemcpyAsync(&hc[0], ctx->dimer_off.as<long long>() + n2, sizeof(long long),
                                   hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            if (int e = grow2(hc[0])) return e;
        }
        launch_dimer_pass(st, n2, ctx->tb_idx.as<int>(), pos, g.g, g.cell_of.as<int>(), g.cell_start.as<int>(),
                          g.members.as<int>(), nullptr, ctx->dimer_off.as<long long>(), ctx->dimers.as<int2>(),
                          ctx->dimer_cap);
    }
    if (want3) {
        CellGrid& g = ctx->grid3;
        if (int e = bin_cells(ctx, st, g, n3, ctx->th_idx.as<int>())) return e;
        launch_fill_zero(st, ctx->lo_count.p, sizeof(int) * (n3 + 1));
        launch_lower_pass(st, n3, ctx->th_idx.as<int>(), pos, g.g, g.cell_of.as<int>(), g.cell_start.as<int>(),
                          g.members.as<int>(), ctx->lo_count.as<int>(), nullptr, nullptr, 0);
        if (int e = scan_i32(ctx, ctx->lo_count.as<int>(), ctx->lo_start.as<int>(), n3 + 1, st)) return e;
        if (ctx->lower_cap == 0) {
            HIP_TRY(hipMemcpyAsync(&hc[1], ctx->lo_start.as<int>() + n3, sizeof(int), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            const int nl = (int)(reinterpret_cast<int*>(&hc[1])[0]);
            HIP_TRY(ctx->lower.ensure(sizeof(int) * (nl + nl / 4 + 64)));
            ctx->lower_cap = nl + nl / 4 + 64;
        }
        launch_lower_pass(st, n3, ctx->th_idx.as<int>(), pos, g.g, g.cell_of.as<int>(), g.cell_start.as<int>(),
                          g.members.as<int>(), nullptr, ctx->lo_start.as<int>(), ctx->lower.as<int>(), ctx->lower_cap);
        launch_fill_zero(st, ctx->trip_count.p, sizeof(long long) * (n3 + 1));
        launch_trimer_pass(st, n3, ctx->th_idx.as<int>(), pos, g.g, ctx->lo_start.as<int>(), ctx->lower.as<int>(),
                           ctx->full_chain_list, ctx->trip_count.as<long long>(), nullptr, nullptr, ctx->lower_cap, 0);
        if (int e = scan_i64(ctx, ctx->trip_count.as<long long>(), ctx->trip_off.as<long long>(), n3 + 1, st)) return e;
        if (ctx->trip_cap == 0) {
            HIP_TRY(hipMemcpyAsync(&hc[2], ctx->trip_off.as<long long>() + n3, sizeof(long long), hipMemcpyDeviceToHost,
                                   st));
            HIP_TRY(hipStreamSynchronize(st));
            const long long nt = hc[2];
            HIP_TRY(ctx->trips.ensure(sizeof(int) * 3 * (nt + nt / 4 + 64)));
            ctx->trip_cap = nt + nt / 4 + 64;
        }
        launch_trimer_pass(st, n3, ctx->th_idx.as<int>(), pos, g.g, ctx->lo_start.as<int>(), ctx->lower.as<int>(),
                           ctx->full_chain_list, nullptr, ctx->trip_off.as<long long>(), ctx->trips.as<int>(),
                           ctx->lower_cap, ctx->trip_cap);
    }
    return MBPOL_OK;
}

// After an evaluation: the list lengths (read back with the energies into
// h_counts[16..18] by results_to_host) against the capacities; when one did not fit, the lists
// grow and `redo` asks for the evaluation to be run again.
static int list_check(mbpol_context* ctx, bool& redo) {
    const mbpol_system& s = ctx->sys;
    const long long* hc = static_cast<const long long*>(ctx->h_counts);
    redo = false;
    if (ctx->lists_built & 1) {
        const long long nd = hc[16];
        if (nd > ctx->dimer_cap) {
            HIP_TRY(ctx->dimers.ensure(sizeof(int2) * (nd + nd / 4 + 64)));
            ctx->dimer_cap = nd + nd / 4 + 64;
            redo = true;
        }
        ctx->stats.dimers = nd;
    }
    if (ctx->lists_built & 2) {
        const long long nl = hc[17], nt = hc[18];
        if (nl > ctx->lower_cap) {
            HIP_TRY(ctx->lower.ensure(sizeof(int) * (nl + nl / 4 + 64)));
            ctx->lower_cap = (int)(nl + nl / 4 + 64);
            redo = true;  // the trimer list could not be built: its length is unknown
        } else if (nt > ctx->trip_cap) {
            HIP_TRY(ctx->trips.ensure(sizeof(int) * 3 * (nt + nt / 4 + 64)));
            ctx->trip_cap = nt + nt / 4 + 64;
            redo = true;
        }
        ctx->stats.trimers = nt;
    }
    (void)s;
    return MBPOL_OK;
}

// device -> host, in one copy: the energy sums into h_counts[8..15] and the
// three list lengths into h_counts[16..18] (gathered by one kernel)
static int results_to_host(mbpol_context* ctx, hipStream_t st) {
    const mbpol_system& s = ctx->sys;
    const int n2 = (int)s.tb_idx.size() / 3, n3 = (int)s.th_idx.size() / 3;
    const bool l2 = ctx->lists_built & 1, l3 = ctx->lists_built & 2;
    u64* stage = ctx->results_stage.as<u64>();
    launch_collect_results(st, ctx->energy_fx.as<u64>(), l2 ? ctx->dimer_off.as<long long>() + n2 : nullptr,
                           l3 ? ctx->lo_start.as<int>() + n3 : nullptr, l3 ? ctx->trip_off.as<long long>() + n3 : nullptr,
                           stage);
    HIP_TRY(hipMemcpyAsync(static_cast<u64*>(ctx->h_counts) + 8, stage, sizeof(u64) * 11, hipMemcpyDeviceToHost, st));
    return MBPOL_OK;
}

// 1-, 2- and 3-body terms and the dispersion, enqueued on stream sr
static int short_range(mbpol_context* ctx, hipStream_t sr, uint32_t comp) {
    const mbpol_system& s = ctx->sys;
    u64* F = ctx->force_fx.as<u64>();
    u64* E = ctx->energy_fx.as<u64>();
    const double* pos = ctx->pos.as<double>();
    const d3 box{s.box[0], s.box[1], s.box[2]};
    const bool do1 = (comp & MBPOL_ONEBODY) && !s.ob_idx.empty();
    const bool do2 = (comp & MBPOL_TWOBODY) && !s.tb_idx.empty();
    const bool do3 = (comp & MBPOL_THREEBODY) && !s.th_idx.empty();
    // independent groups: per-unit energies, summed per group after each term
    const int G = ctx->ngroups;
    u64* EG = ctx->g_energy.as<u64>();
    if (do1 && !ctx->onebody_with_charges) {
        timer_start(ctx, T_ONEBODY);
        launch_water_terms(sr, (int)s.ob_idx.size() / 3, true, false, ctx->ob_idx.as<int>(), pos,
                           s.ob_method == MBPOL_ONEBODY_PERIODIC, box, F, E + 0, G ? ctx->g_eu1.as<double>() : nullptr,
                           nullptr, nullptr);
        if (G) launch_group_sums(sr, G, ctx->g_wstart.as<int>(), 1, nullptr, ctx->g_eu1.as<double>(), EG, 0);
        timer_stop(ctx, T_ONEBODY);
    }
    if (do2 || do3) {
        timer_start(ctx, T_LISTS);
        if (int e = build_lists(ctx, sr, do2, do3)) return e;
        timer_stop(ctx, T_LISTS);
    }
    ctx->lists_built = (do2 ? 1 : 0) | (do3 ? 2 : 0);
    if (G && do2) HIP_TRY(ctx->g_eu2.ensure(sizeof(double) * std::max<long long>(ctx->dimer_cap, 1)));
    if (G && do3) HIP_TRY(ctx->g_eu3.ensure(sizeof(double) * std::max<long long>(ctx->trip_cap, 1)));
    double* eu2 = G ? ctx->g_eu2.as<double>() : nullptr;
    double* eu3 = G ? ctx->g_eu3.as<double>() : nullptr;
    // The polynomials run in FP64 unless FP32 was asked for explicitly: their
    // terms cancel to ~1e-3 of their magnitudes, so FP32 evaluation loses three
    // digits of the 2- and 3-body energies.
    const bool dp = ctx->precision != MBPOL_PRECISION_SINGLE_PIP;
    auto run2 = [&]() -> int {
        if (do2) {
            const int per = s.tb_method == MBPOL_CUTOFF_PERIODIC;
            const KEvents ev = timer_kernel(ctx, T_TWOBODY);
            const long long* nd = ctx->dimer_off.as<long long>() + (int)s.tb_idx.size() / 3;  // list length (device)
            // the polynomial through the dimer's relabelling group, or term by term (MBPOL_HIP_2B_KERNEL=terms)
            auto launch2 = dp ? (ctx->pip2_kernel == 1 ? launch_twobody_f64 : launch_twobody_orbit_f64)
                              : (ctx->pip2_kernel == 1 ? launch_twobody_f32 : launch_twobody_orbit_f32);
            launch2(sr, ctx->dimer_cap, nd, ctx->dimers.as<int2>(), ctx->tb_idx.as<int>(), pos, per, box, F, E + 1, ev,
                    eu2, ctx->pip_blocks);
            timer_kernel_done(ctx, T_TWOBODY, ev);
            // dimers are listed by their first molecule, so a group's dimers are consecutive
            if (G)
                launch_group_sums(sr, G, ctx->g_wstart.as<int>(), 0, ctx->dimer_off.as<long long>(), eu2, EG, 1,
                                  ctx->dimer_cap);
        }
        return MBPOL_OK;
    };
    auto run3 = [&]() -> int {
        if (do3) {
            const int per = s.th_method == MBPOL_CUTOFF_PERIODIC;
            const KEvents ev = timer_kernel(ctx, T_THREEBODY);
            const long long nt = ctx->trip_cap;
            const long long* ntd = ctx->trip_off.as<long long>() + (int)s.th_idx.size() / 3;  // list length (device)
            const int* tr = ctx->trips.as<int>();
            const int* ti = ctx->th_idx.as<int>();
            const int cap = ctx->pip3_blocks;
            if (!dp)
                launch_threebody_orbit_f32(sr, nt, ntd, tr, ti, pos, per, box, F, E + 2, ev, eu3, cap);
            else if (ctx->pip3_kernel == 1)
                launch_threebody_f64(sr, nt, ntd, tr, ti, pos, per, box, F, E + 2, ev, eu3, cap);
            else
                launch_threebody_orbit_f64(sr, nt, ntd, tr, ti, pos, per, box, F, E + 2, ev, eu3, cap);
            timer_kernel_done(ctx, T_THREEBODY, ev);
            if (G)
                launch_group_sums(sr, G, ctx->g_wstart.as<int>(), 0, ctx->trip_off.as<long long>(), eu3, EG, 2,
                                  ctx->trip_cap);
        }
        return MBPOL_OK;
    };
    auto rund = [&]() -> int {
        if ((comp & MBPOL_DISPERSION) && ctx->disp) {
            timer_start(ctx, T_DISPERSION);
            if (int e = disp_evaluate(ctx, sr, E + 3)) return e;
            if (G)
                if (int e = disp_group_sums(ctx, sr, EG)) return e;
            timer_stop(ctx, T_DISPERSION);
        }
        return MBPOL_OK;
    };
    for (const char* o = ctx->sr_order; *o; ++o) {
        const int e = *o == '2' ? run2() : *o == '3' ? run3() : *o == 'd' ? rund() : MBPOL_OK;
        if (e) return e;
    }
    return MBPOL_OK;
}

// Everything of one evaluation, enqueued on the context's streams, up to the
// copies to the host of the energies, the list lengths and (forces_host_out, a
// pinned host buffer) the forces.  fixed_steps > 0 asks for the dipole solve of
// that many field evaluations with no host read: the form a HIP graph captures.
static int enqueue_evaluation(mbpol_context* ctx, double* forces_dev_out, uint32_t comp, double* forces_host_out,
                              int fixed_steps) {
    const mbpol_system& s = ctx->sys;
    hipStream_t st = ctx->stream;
    u64* F = ctx->force_fx.as<u64>();
    u64* E = ctx->energy_fx.as<u64>();
    double* pos = ctx->pos.as<double>();
    {
        FillList L(st);
        L.add(F, sizeof(u64) * 3 * s.n);
        L.add(E, sizeof(u64) * 8);
        if (ctx->ngroups) L.add(ctx->g_energy.p, sizeof(u64) * 8 * ctx->ngroups);
        L.launch();
    }
    ctx->lists_built = 0;
    launch_place_vsites(st, (int)s.vs_idx.size() / 4, ctx->vs_idx.as<int>(), ctx->vs_w.as<double>(), pos);
    // the 1-body term runs with the charges when both are computed for the same waters
    ctx->onebody_with_charges = ctx->water_fused && (comp & MBPOL_ONEBODY) && (comp & MBPOL_ELECTROSTATICS) &&
                                ctx->elec && !s.ob_idx.empty();
    // The short-range terms go to a second stream, forked from and joined to the
    // electrostatics' stream by events, so that the two overlap; no host wait sits
    // in either (list lengths and row counts stay on the device and are checked
    // after the evaluation).  With timing on, everything runs on one stream, so
    // that every timed span is attributable; the first evaluation also runs the
    // launch-shape tuning, which times kernels, and sizes the lists: unforked.
    // The forked short-range work is enqueued after the electrostatics set-up
    // (whose kernels keep the GPU busy while the host enqueues it) and before
    // the dipole solve, whose host reads would otherwise delay it.  The fork
    // point on the GPU is either right after the virtual sites are placed
    // (fork_early: the short-range kernels also run beside the set-up's many
    // short kernels) or after the set-up; or, third, right after the virtual
    // sites with the short-range work enqueued at once by the side enqueuer's
    // thread while this one enqueues the electrostatics (not under capture);
    // the tuning picks one.
    const bool fork = !ctx->timer.enabled && ctx->evaluated;
    const bool elec = (comp & MBPOL_ELECTROSTATICS) && ctx->elec;
    const bool threaded = fork && elec && ctx->fork_early == 2 && fixed_steps == 0 && !mbpol::g_capturing;
    if (fork && ctx->fork_early) HIP_TRY(hipEventRecord(ctx->ev_fork, st));
    int rc = MBPOL_OK;
    bool forked = false;
    auto fork_short_range = [&]() -> int {
        if (!ctx->fork_early) HIP_TRY(hipEventRecord(ctx->ev_fork, st));
        HIP_TRY(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
        forked = true;
        const int e = short_range(ctx, ctx->side, comp);
        // recorded also after a failure, so that the join below completes a capture
        HIP_TRY(hipEventRecord(ctx->ev_join, ctx->side));
        return e;
    };
    if (!fork) {
        rc = short_range(ctx, st, comp);
        if (!rc && elec) rc = elec_evaluate(ctx, E + 4, nullptr, fixed_steps);
    } else if (threaded) {
        if (int e = side_post(ctx, fork_short_range)) return e;
        rc = elec_evaluate(ctx, E + 4, nullptr, fixed_steps);
        // always waited for: the job refers to this frame
        const int e = side_wait(ctx);
        if (!rc) rc = e;
    } else if (elec) {
        rc = elec_evaluate(ctx, E + 4, fork_short_range, fixed_steps);
        if (rc && !forked) {
            // the solve failed before the short-range work was enqueued: enqueue it now, so that
            // a re-run finds no half-built lists (and a capture joins its streams)
            const int e = fork_short_range();
            (void)e;
        }
    } else {
        rc = fork_short_range();
    }
    if (forked) HIP_TRY(hipStreamWaitEvent(st, ctx->ev_join, 0));
    if (rc) return rc;
    double* fout = forces_dev_out ? forces_dev_out : ctx->force.as<double>();
    launch_forces_to_double(st, 3 * s.n, F, fout);
    launch_fold_vsites(st, (int)s.vs_idx.size() / 4, ctx->vs_idx.as<int>(), ctx->vs_w.as<double>(), fout);
    if (int e = results_to_host(ctx, st)) return e;
    if (forces_host_out)
        HIP_TRY(hipMemcpyAsync(forces_host_out, fout, sizeof(double) * 3 * s.n, hipMemcpyDeviceToHost, st));
    return MBPOL_OK;
}

// the five component energies and their total, from the fixed-point sums read back
static void energies_from_host(const mbpol_context* ctx, double* energies_out) {
    if (!energies_out) return;
    const u64* he = static_cast<const u64*>(ctx->h_counts) + 8;
    double tot = 0;
    for (int k = 0; k < 5; ++k) {
        energies_out[k] = (double)(long long)he[k] * kFixInvHost;
        tot += energies_out[k];
    }
    energies_out[5] = tot;
}

// Evaluation-level tuning of the overlap: the fork of the short-range stream
// and the short-range grid caps.  The 2-body, 3-body and dispersion kernels run
// beside the electrostatics, so what counts is the whole overlapped evaluation,
// not a kernel alone: on the first forked evaluation each knob in turn -- the
// fork (early, after the electrostatics set-up, or early with the short-range
// work enqueued by a host thread of its own), then each cap (waves per SIMD;
// 0 = one block per 64 units) -- has its candidates timed on trial evaluations
// of the caller's configuration, from the first launch to the final
// synchronisation on the host.  The candidates take turns (three rounds of
// three evaluations each), so that a drift of the clocks falls on all of them
// alike; a candidate replaces the current setting when its median is 1.5%
// faster.  Every candidate leaves the results bit for bit the same
// (fixed-point force and energy sums, per-unit energies summed in list order),
// so the tuning is invisible but in time.  The trial solves start from
// alpha E_fixed and the previous dipoles are put back afterwards.
// MBPOL_HIP_TUNE=0 turns it off; a setting fixed by its environment variable
// (MBPOL_HIP_FORK=early|late|thread, MBPOL_HIP_*_WAVES) is not tuned.
static int evaluate_eager(mbpol_context* ctx, double* forces_dev_out, double* energies_out, uint32_t comp,
                          double* forces_host_out);

static int tune_overlap(mbpol_context* ctx, uint32_t comp) {
    ctx->tuned = true;
    if (const char* env = std::getenv("MBPOL_HIP_TUNE"))
        if (std::atoi(env) == 0) return MBPOL_OK;
    const mbpol_system& s = ctx->sys;
    const bool elec = (comp & MBPOL_ELECTROSTATICS) && ctx->elec;
    if (elec && !elec_tune_ok(ctx)) return MBPOL_OK;
    struct Knob {
        const char* env;
        bool on;
        int* waves;
        int* blocks;  // null: the value itself is the setting
        std::vector<int> cand;
    } knobs[] = {
        {"MBPOL_HIP_FORK", true, &ctx->fork_early, nullptr, {2, 1, 0}},
        {"MBPOL_HIP_3B_WAVES", (comp & MBPOL_THREEBODY) && !s.th_idx.empty(), &ctx->pip3_waves, &ctx->pip3_blocks,
         {1, 0, 2}},
        {"MBPOL_HIP_DISP_WAVES", (comp & MBPOL_DISPERSION) && ctx->disp, &ctx->disp_waves, &ctx->disp_blocks,
         {2, 1, 0}},
        {"MBPOL_HIP_PIP_WAVES", (comp & MBPOL_TWOBODY) && !s.tb_idx.empty(), &ctx->pip_waves, &ctx->pip_blocks,
         {1, 2, 0}},
    };
    if (elec)
        if (int e = elec_tune_hold(ctx, true)) return e;
    hipStream_t st = ctx->stream;
    bool ok = true;
    auto set = [&](Knob& k, int w) {
        *k.waves = w;
        if (k.blocks) *k.blocks = w * 4 * ctx->cus;
    };
    // host time of `count` trial evaluations appended to t; false when one failed or did not fit
    auto measure = [&](int count, std::vector<double>& t) -> bool {
        for (int r = 0; r < count; ++r) {
            const auto t0 = std::chrono::steady_clock::now();
            const int rc = enqueue_evaluation(ctx, nullptr, comp, nullptr, 0);
            if (hipStreamSynchronize(st) != hipSuccess || rc != MBPOL_OK) return false;
            t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
            ++ctx->tune_evaluations;
            bool redo = false, redo_lists = false;
            if (elec && elec_rows_check(ctx, redo) != MBPOL_OK) return false;
            if (list_check(ctx, redo_lists) != MBPOL_OK || redo || redo_lists) return false;
        }
        return true;
    };
    auto median = [](std::vector<double> v) {
        std::sort(v.begin(), v.end());
        return v[v.size() / 2];
    };
    for (Knob& k : knobs) {
        if (!ok || !k.on || std::getenv(k.env)) continue;
        const int start = *k.waves;
        std::vector<int> cand{start};
        for (int w : k.cand)
            if (w != start) cand.push_back(w);
        std::vector<std::vector<double>> t(cand.size());
        for (int round = 0; ok && round < 3; ++round)
            for (size_t c = 0; ok && c < cand.size(); ++c) {
                set(k, cand[c]);
                ok = measure(3, t[c]);
            }
        int best = start;
        if (ok) {
            double best_ms = median(t[0]);
            for (size_t c = 1; c < cand.size(); ++c) {
                const double ms = median(t[c]);
                if (ms < 0.985 * best_ms) {
                    best_ms = ms;
                    best = cand[c];
                }
            }
        }
        set(k, best);
    }
    (void)hipGetLastError();
    HIP_TRY(hipStreamSynchronize(st));
    if (elec)
        if (int e = elec_tune_hold(ctx, false)) return e;
    return MBPOL_OK;
}

static int evaluate_eager(mbpol_context* ctx, double* forces_dev_out, double* energies_out, uint32_t comp,
                          double* forces_host_out) {
    hipStream_t st = ctx->stream;
    const bool fork = !ctx->timer.enabled && ctx->evaluated;
    if (fork && !ctx->tuned)
        if (int e = tune_overlap(ctx, comp)) return e;
    const int rc = enqueue_evaluation(ctx, forces_dev_out, comp, forces_host_out, 0);
    if (rc) {
        // the short-range kernels may still be adding into the accumulators and
        // reading the positions, which a re-run (or the next call) clears and overwrites
        HIP_TRY(hipStreamSynchronize(st));
        // a solve that failed on rows or lists that overflowed is run again with ones that fit
        bool redo = false, redo_lists = false;
        if (ctx->elec)
            if (int e = elec_rows_check(ctx, redo)) return e;
        if (int e = results_to_host(ctx, st)) return e;
        HIP_TRY(hipStreamSynchronize(st));
        if (int e = list_check(ctx, redo_lists)) return e;
        if (redo || redo_lists) return evaluate_eager(ctx, forces_dev_out, energies_out, comp, forces_host_out);
        return rc;
    }
    HIP_TRY(hipStreamSynchronize(st));
    {
        // rows or lists that did not fit: grown, and the evaluation runs again (once:
        // the capacities then hold the lengths just measured)
        bool redo = false, redo_lists = false;
        if (ctx->elec)
            if (int e = elec_rows_check(ctx, redo)) return e;
        if (int e = list_check(ctx, redo_lists)) return e;
        if (redo || redo_lists) return evaluate_eager(ctx, forces_dev_out, energies_out, comp, forces_host_out);
    }
    ctx->evaluated = true;
    if (fork) ctx->fork_ready = true;
    energies_from_host(ctx, energies_out);
    return MBPOL_OK;
}

// The host state a captured evaluation was recorded from: the arguments, the
// parameter generation, the list and row capacities (their buffers' addresses
// follow them), the solve's length and the electrostatics' guess start.
static void graph_key(const mbpol_context* ctx, double* forces_dev_out, uint32_t comp, double* forces_host_out,
                      long long key[16]) {
    std::memset(key, 0, sizeof(long long) * 16);
    key[0] = comp;
    key[1] = (long long)(intptr_t)forces_dev_out;
    key[2] = (long long)(intptr_t)forces_host_out;
    key[3] = ctx->graph_gen;
    key[4] = ctx->dimer_cap;
    key[5] = ctx->lower_cap;
    key[6] = ctx->trip_cap;
    key[7] = ctx->graph_steps;
    key[8] = (long long)(intptr_t)ctx->pos.p;
    if ((comp & MBPOL_ELECTROSTATICS) && ctx->elec) elec_graph_key(ctx, key + 9);
}

static int capture_evaluation(mbpol_context* ctx, double* forces_dev_out, uint32_t comp, double* forces_host_out) {
    hipStream_t st = ctx->stream;
    HIP_TRY(hipStreamSynchronize(st));
    if (ctx->graph_exec) {
        HIP_TRY(hipGraphExecDestroy(ctx->graph_exec));
        ctx->graph_exec = nullptr;
    }
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    mbpol::g_capturing = true;
    const int rc = enqueue_evaluation(ctx, forces_dev_out, comp, forces_host_out, ctx->graph_steps);
    mbpol::g_capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &g);
    if (rc || ec != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return rc ? rc : set_error(MBPOL_ERR_HIP, "capture of the evaluation failed: %s", hipGetErrorString(ec));
    }
    const hipError_t ei = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        ctx->graph_exec = nullptr;
        (void)hipGetLastError();
        return set_error(MBPOL_ERR_HIP, "instantiation of the evaluation graph failed: %s", hipGetErrorString(ei));
    }
    ctx->graph_lists_built = ctx->lists_built;
    ++ctx->graph_captures;
    return MBPOL_OK;
}

// Graph mode: the evaluation is captured once (after an eager, forked one has
// sized every buffer) and replayed while the host state it was recorded from
// holds.  After a replay the host runs the eager path's checks: a list or row
// that overflowed, or a dipole solve that needed more field evaluations than
// were captured, has the evaluation run again eagerly (and captured anew on the
// next call).  `used` is false when the evaluation is not one a graph can hold
// (then nothing has been enqueued).
// field evaluations of the captured solve: the last solve's + 1, or MBPOL_HIP_GRAPH_STEPS
static int graph_steps_for(const mbpol_context* ctx) {
    if (const char* env = std::getenv("MBPOL_HIP_GRAPH_STEPS")) {
        const int v = std::atoi(env);
        if (v > 0) return std::min(v, ctx->sys.el_max_iter + 1);
    }
    return elec_graph_steps(ctx, 1);
}

static int evaluate_graph(mbpol_context* ctx, double* forces_dev_out, double* energies_out, uint32_t comp,
                          double* forces_host_out, bool& used) {
    used = false;
    hipStream_t st = ctx->stream;
    const bool elec = (comp & MBPOL_ELECTROSTATICS) && ctx->elec;
    if (elec && !elec_graph_ok(ctx)) return MBPOL_OK;
    if (elec && ctx->graph_steps == 0) ctx->graph_steps = graph_steps_for(ctx);
    long long key[16];
    graph_key(ctx, forces_dev_out, comp, forces_host_out, key);
    if (!ctx->graph_exec || std::memcmp(key, ctx->graph_key, sizeof(key)) != 0) {
        if (std::memcmp(key, ctx->graph_bad_key, sizeof(key)) == 0) return MBPOL_OK;
        if (capture_evaluation(ctx, forces_dev_out, comp, forces_host_out) != MBPOL_OK) {
            // eager from here on for this state; the stream must be idle before it is reused
            std::memcpy(ctx->graph_bad_key, key, sizeof(key));
            std::memset(ctx->graph_key, 0, sizeof(key));
            HIP_TRY(hipStreamSynchronize(st));
            return MBPOL_OK;
        }
        std::memcpy(ctx->graph_key, key, sizeof(key));
    }
    used = true;
    HIP_TRY(hipGraphLaunch(ctx->graph_exec, st));
    HIP_TRY(hipStreamSynchronize(st));
    ++ctx->graph_replays;
    ctx->lists_built = ctx->graph_lists_built;
    bool redo = false, redo_lists = false, more = false;
    int rc = MBPOL_OK;
    if (elec) {
        rc = elec_graph_finish(ctx, more);
        if (int e = elec_rows_check(ctx, redo)) return e;
    }
    if (int e = list_check(ctx, redo_lists)) return e;
    if (redo || redo_lists || more) {
        ++ctx->graph_fallbacks;
        const int r = evaluate_eager(ctx, forces_dev_out, energies_out, comp, forces_host_out);
        if (more && elec) ctx->graph_steps = graph_steps_for(ctx);
        return r;
    }
    if (rc) return rc;
    energies_from_host(ctx, energies_out);
    return MBPOL_OK;
}

static int evaluate(mbpol_context* ctx, double* forces_dev_out, double* energies_out, uint32_t comp,
                    double* forces_host_out = nullptr) {
    if (ctx->graph_on && ctx->fork_ready && !ctx->timer.enabled) {
        bool used = false;
        const int rc = evaluate_graph(ctx, forces_dev_out, energies_out, comp, forces_host_out, used);
        if (used) return rc;
        if (rc) return rc;
    }
    return evaluate_eager(ctx, forces_dev_out, energies_out, comp, forces_host_out);
}

extern "C" {

int mbpol_context_evaluate(mbpol_context* ctx, const double* positions, double* forces_out, double* energies_out,
                           uint32_t components) {
    if (!ctx || !positions) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    // through the context's pinned buffers: copies from and to pageable memory
    // would be staged by the runtime in small synchronous pieces
    const size_t bytes = sizeof(double) * 3 * ctx->sys.n;
    HIP_TRY(hipStreamSynchronize(ctx->stream));  // no earlier copy still reads the staging buffer
    std::memcpy(ctx->h_pos, positions, bytes);
    HIP_TRY(hipMemcpyAsync(ctx->pos.p, ctx->h_pos, bytes, hipMemcpyHostToDevice, ctx->stream));
    timer_start(ctx, T_EVALUATE);
    if (int e = evaluate(ctx, nullptr, energies_out, components, forces_out ? ctx->h_force : nullptr)) return e;
    timer_stop(ctx, T_EVALUATE);
    if (forces_out) std::memcpy(forces_out, ctx->h_force, bytes);
    return MBPOL_OK;
}

int mbpol_context_evaluate_device(mbpol_context* ctx, const double* d_positions, double* d_forces, double* energies_out,
                                  uint32_t components) {
    if (!ctx || !d_positions) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    if (d_positions != ctx->pos.p)
        HIP_TRY(hipMemcpyAsync(ctx->pos.p, d_positions, sizeof(double) * 3 * ctx->sys.n, hipMemcpyDeviceToDevice,
                               ctx->stream));
    timer_start(ctx, T_EVALUATE);
    if (int e = evaluate(ctx, d_forces, energies_out, components)) return e;
    timer_stop(ctx, T_EVALUATE);
    return MBPOL_OK;
}

int mbpol_context_upload_positions(mbpol_context* ctx, const double* positions) {
    if (!ctx || !positions) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(copy_on(ctx->stream, ctx->pos.p, positions, sizeof(double) * 3 * ctx->sys.n));
    // keep an untouched copy: evaluate() rewrites the virtual-site rows in place
    HIP_TRY(ctx->pos_resident.ensure(sizeof(double) * 3 * ctx->sys.n));
    HIP_TRY(copy_on(ctx->stream, ctx->pos_resident.p, ctx->pos.p, sizeof(double) * 3 * ctx->sys.n));
    return MBPOL_OK;
}

int mbpol_context_evaluate_resident(mbpol_context* ctx, double* energies_out, uint32_t components) {
    if (!ctx || !ctx->pos_resident.p) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "no resident positions");
    return mbpol_context_evaluate_device(ctx, ctx->pos_resident.as<double>(), nullptr, energies_out, components);
}

int mbpol_context_download_forces(mbpol_context* ctx, double* forces_out) {
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(copy_on(ctx->stream, forces_out, ctx->force.p, sizeof(double) * 3 * ctx->sys.n));
    return MBPOL_OK;
}

int mbpol_context_set_timing(mbpol_context* ctx, int32_t enabled) {
    ++ctx->graph_gen;
    timer_collect(ctx);
    ctx->timer.enabled = enabled != 0;
    for (int c = 0; c < T_NCAT; ++c) {
        ctx->timer.ms[c] = 0;
        ctx->timer.count[c] = 0;
    }
    return MBPOL_OK;
}

int mbpol_context_get_timing(mbpol_context* ctx, const char* name, double* total_ms, int64_t* count) {
    static const char* names[T_NCAT] = {"evaluate",   "onebody",    "lists",        "twobody",
                                        "threebody",  "dispersion", "elec_setup",   "elec_iteration",
                                        "elec_energy", "pme_spread", "fft_z_r2c",   "fft_y",
                                        "fft_x_conv", "fft_z_c2r",  "elec_iterate", "fft_plane_fwd",
                                        "fft_plane_inv", "spread_fft_z"};
    timer_collect(ctx);
    for (int c = 0; c < T_NCAT; ++c)
        if (std::strcmp(name, names[c]) == 0) {
            *total_ms = ctx->timer.ms[c];
            *count = ctx->timer.count[c];
            return MBPOL_OK;
        }
    return set_error(MBPOL_ERR_INVALID_ARGUMENT, "unknown timer '%s'", name);
}

int mbpol_context_get_group_energies(const mbpol_context* ctx, double* energies) {
    if (!ctx || !energies) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->ngroups) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "the system has no groups");
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<u64> h((size_t)8 * ctx->ngroups);
    HIP_TRY(copy_on(ctx->stream, h.data(), ctx->g_energy.p, sizeof(u64) * h.size()));
    for (int g = 0; g < ctx->ngroups; ++g) {
        double tot = 0;
        for (int k = 0; k < 5; ++k) {
            energies[6 * g + k] = (double)(long long)h[8 * (size_t)g + k] * kFixInvHost;
            tot += energies[6 * g + k];
        }
        energies[6 * g + 5] = tot;
    }
    return MBPOL_OK;
}

int mbpol_context_get_stats(const mbpol_context* ctx, mbpol_eval_stats* out) {
    out->iterations = ctx->stats.iterations;
    out->epsilon = ctx->stats.epsilon;
    out->converged = ctx->stats.converged;
    out->num_dimers = ctx->stats.dimers;
    out->num_trimers = ctx->stats.trimers;
    out->num_elec_pairs = ctx->stats.elec_pairs;
    long long dp = 0;
    if (ctx->disp) {
        (void)hipSetDevice(ctx->device);
        if (int e = disp_pair_count(ctx, &dp)) return e;
    }
    out->num_disp_pairs = dp;
    out->num_pol_pairs = ctx->stats.pol_pairs;
    return MBPOL_OK;
}

int mbpol_context_get_dimer_list(const mbpol_context* ctx, int32_t* pairs, int64_t capacity, int64_t* count) {
    *count = ctx->stats.dimers;
    if (pairs && capacity > 0 && ctx->stats.dimers > 0) {
        const long long m = std::min<long long>(capacity, ctx->stats.dimers);
        HIP_TRY(copy_on(ctx->stream, pairs, ctx->dimers.p, sizeof(int2) * m));
    }
    return MBPOL_OK;
}

int mbpol_context_get_trimer_list(const mbpol_context* ctx, int32_t* trip, int64_t capacity, int64_t* count) {
    *count = ctx->stats.trimers;
    if (trip && capacity > 0 && ctx->stats.trimers > 0) {
        const long long m = std::min<long long>(capacity, ctx->stats.trimers);
        HIP_TRY(copy_on(ctx->stream, trip, ctx->trips.p, sizeof(int) * 3 * m));
    }
    return MBPOL_OK;
}

}  // extern "C"
