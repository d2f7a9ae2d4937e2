// Host driver of the electrostatics: PME parameters, B-spline moduli, hipFFT
// plans, the induced-dipole iteration and the energy/force pass.
// Follows MBPolReferenceKernels.cpp:141-259 (set-up) and
// MBPolReferenceElectrostaticsForce.cpp:557-647, :857-921, :2718-2777.
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mbpol_hip.h"
#include "elec.hpp"
#include "elec_kernels.h"
#include "fft_kernels.h"
#include "mbpol_params.h"

namespace mbpol {

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t _e = (expr);                                                                         \
        if (_e != hipSuccess) return set_error(MBPOL_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)
#define FFT_TRY(expr)                                                                        \
    do {                                                                                     \
        hipfftResult _r = (expr);                                                            \
        if (_r != HIPFFT_SUCCESS) return set_error(MBPOL_ERR_HIP, "%s: hipFFT error %d", #expr, (int)_r); \
    } while (0)

// ---------------------------------------------------------------------------
// Potential at points without allocation (mbpol_context_get_electrostatic_potential_noalloc):
// the points and the results are staged through `scratch`, a fixed device array
// of this library (2 MiB; every GPU holds its own copy until the process ends).
// Its device address is looked up once per device, with that device current,
// when the first context with electrostatics is created on it.  Calls that
// overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
namespace {
constexpr int kScratchPoints = 1 << 16;                          // at most this many points per call
static __device__ __align__(256) double scratch[4 * kScratchPoints];  // 3 coordinates per point, then the potentials
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// the calling thread's current device must be `device`
static int scratch_address(int device, double** out) {
    if (device < 0 || device >= kMaxDevices) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[device];
    return MBPOL_OK;
}

struct Elec {
    ElecParams P{};
    int n = 0, nw = 0, npol = 0;
    int guess_mode = MBPOL_GUESS_ALPHA_E;
    int solver = MBPOL_SOLVER_SOR;
    bool have_previous = false;
    // parameters; polidx lists the sites of non-zero polarisability
    DevBuf q0, q, damp6, pol, mol, type, dq, polidx;  // damp6: damping^(1/6)
    std::vector<unsigned char> h_pol;  // host copy of (polarisability != 0)
    // cells for the neighbour rows
    DevBuf cell_of, cell_count, cell_start, cell_fill, members, unsorted;
    DevBuf cpos;  // positions and polarisable flags in cell order
    int3 nc{1, 1, 1};
    int reach = 1, use_cells = 0;
    int site_lanes = 0;  // lanes per site of the row build: 0 until tuned on the first evaluation
    // neighbour rows (ELL) and induced-field prefactors
    DevBuf nbP, cntP, nbQ, cntQ, pf, maxcnt;
    int capP = 0, capQ = 0;
    // fields and dipoles: the iteration alternates between mu_a and mu_b
    DevBuf efix, efix_pol, mu_a, mu_b, mu_prev, mu_user, partial, eps;
    // history for the learned guess of the previous-step kind (kept while history > 0): alpha E_fixed
    // of the last evaluation, and the dipoles and alpha E_fixed of the one before it
    int history = 0;
    // state held while the evaluation-level tuning runs (elec_tune_hold)
    DevBuf mu_held;
    bool held_have_previous = false;
    int held_last_iters = 0, held_guess = MBPOL_GUESS_ALPHA_E;
    bool have_previous2 = false;
    DevBuf efix_prev, mu_prev2, efix_prev2;
    // conjugate-gradient vectors (3 n), block partials and the r.z history
    DevBuf cg_r, cg_z, cg_p, cg_q, cg_partial, cg_rz;
    // block-Jacobi preconditioner: polarisable sites grouped by molecule, and the blocks' inverses
    DevBuf blk_start, blk_site, minv;
    int nblk = 0;
    // independent groups: site ranges, per-group preconditioner blocks and site counts, per-block
    // |alpha r|^2, per-site energies
    int ngroups = 0, cap_grp_pol = 0, cap_grp_other = 0;
    DevBuf g_site_s, g_site_e, g_blk_start, g_nsites, g_blk_ee, g_esite;
    double* mu_cur = nullptr;  // dipoles of the last evaluate
    DevBuf phi_fix, phi_ind, fsite, V, eblk;
    // PME
    DevBuf grid, spec, influence, bx, by, bz;
    BrickGeom B{};
    struct Bins {
        DevBuf of, count, start, fill, members;
    } binq;  // sites with a charge, by grid brick
    DevBuf g0tab, sptab;  // per-site spline table
    DevBuf rowcount, rowstart, rowfill, esite, ecoef;  // dipole-spreading table in grid-point order
    // grid transforms: the LDS transforms of kernels_fft.hip when every axis
    // factors into radices 2..13, else hipFFT
    PmeFft fft{};
    DevBuf tw[3];
    bool lds_fft = false;
    bool plane_fft = false;  // z and y passes fused per x plane
    bool fuse_spread = true;  // the dipole spreading fused with the z pass (single-axis passes only)
    hipfftHandle plan_f = 0, plan_b = 0;
    bool plans = false;
    double* h_eps = nullptr;  // pinned, 128 bytes: [0] epsilon, ints from byte 32, a SolveCtl from byte 64
    DevBuf ctl;               // SolveCtl of the running dipole solve
    int last_iters = 0;       // field evaluations of the previous solve (sizes the first batch of steps)
    // row capacities are checked at the end of the evaluation (rows_pending) once
    // the launch shapes are tuned; after an overflow the next build waits for the check
    bool rows_pending = false, rows_sync = true;
};

// B-spline moduli with the reference's small-modulus fix-up and order-5 zeta
// correction (MBPolReferenceElectrostaticsForce.cpp:1248-1340), host FP64.
static std::vector<double> bspline_moduli(int n) {
    const int order = mbpol_params::el_pme_order;
    double arr[8];
    const double x = 0.0;
    arr[0] = 1.0 - x;
    arr[1] = x;
    for (int k = 2; k < order; ++k) {
        const double den = 1.0 / k;
        arr[k] = x * arr[k - 1] * den;
        for (int i = 1; i < k; ++i) arr[k - i] = ((x + i) * arr[k - i - 1] + ((k - i + 1) - x) * arr[k - i]) * den;
        arr[0] = (1.0 - x) * arr[0] * den;
    }
    std::vector<double> bs(std::max(n, order + 2) + 1, 0.0);
    for (int i = 2; i <= order + 1; ++i) bs[i] = arr[i - 2];
    std::vector<double> mod(n, 0.0);
    const double factor = 2.0 * M_PI / n;
    for (int i = 0; i < n; ++i) {
        double s1 = 0, s2 = 0;
        for (int j = 1; j <= n; ++j) {
            const double ang = factor * i * (j - 1);
            s1 += bs[j] * std::cos(ang);
            s2 += bs[j] * std::sin(ang);
        }
        mod[i] = s1 * s1 + s2 * s2;
    }
    const double tiny = 1.0e-7;
    if (mod[0] < tiny) mod[0] = 0.5 * mod[1];
    for (int i = 1; i < n - 1; ++i)
        if (mod[i] < tiny) mod[i] = 0.5 * (mod[i - 1] + mod[i + 1]);
    if (mod[n - 1] < tiny) mod[n - 1] = 0.5 * mod[n - 2];
    for (int i = 1; i <= n; ++i) {
        int k = i - 1;
        if (i > n / 2) k = k - n;
        double zeta = 1.0;
        if (k != 0) {
            double s1 = 1.0, s2 = 1.0;
            const double f = M_PI * k / n;
            for (int j = 1; j <= 50; ++j) {
                const double a = f / (f + M_PI * j);
                s1 += std::pow(a, order);
                s2 += std::pow(a, 2 * order);
            }
            for (int j = 1; j <= 50; ++j) {
                const double a = f / (f - M_PI * j);
                s1 += std::pow(a, order);
                s2 += std::pow(a, 2 * order);
            }
            zeta = s2 / s1;
        }
        mod[i - 1] *= zeta * zeta;
    }
    return mod;
}

static SiteParams site_params(Elec* E) {
    return SiteParams{E->q.as<double>(), E->damp6.as<double>(), E->pol.as<double>(), E->mol.as<int>(), E->type.as<int>()};
}

// per-site damping^(1/6): the pair's Thole width (damp_i damp_j)^(1/6) is their product
static std::vector<double> sixth_roots(const std::vector<double>& d) {
    std::vector<double> out(d.size());
    for (size_t k = 0; k < d.size(); ++k) out[k] = std::pow(d[k], 1.0 / 6.0);
    return out;
}

template <typename T>
static int up(hipStream_t stream, DevBuf& b, const std::vector<T>& v) {
    HIP_TRY(b.ensure(std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty()) HIP_TRY(copy_on(stream, b.p, v.data(), v.size() * sizeof(T)));
    return MBPOL_OK;
}

// factor n into the radices of the LDS transforms; false if some factor is not one of them
static bool fft_factor(int n, FftAxis& a) {
    a.n = n;
    a.nfac = 0;
    int r = n;
    for (int p : {4, 2, 3, 5, 7, 11, 13})
        while (r % p == 0 && a.nfac < 12) {
            a.fac[a.nfac++] = p;
            r /= p;
        }
    if (r != 1 || n > kFftMaxLength) return false;
    int ns = 1, maxnb = 1;
    for (int s = 0; s < a.nfac; ++s) {
        a.ns[s] = ns;
        a.nb[s] = n / a.fac[s];
        a.step[s] = n / (ns * a.fac[s]);
        a.inv_ns[s] = 1.0f / (float)ns;
        maxnb = std::max(maxnb, a.nb[s]);
        ns *= a.fac[s];
    }
    a.lgP = 0;
    while ((1 << a.lgP) < maxnb && a.lgP < 7) ++a.lgP;
    a.inplace = (1 << a.lgP) >= maxnb;
    return true;
}

// Brick size along one grid axis for the spreading kernel: every brick at least
// 4 points wide and none above 12 (the last one takes the remainder).
static void brick_axis(int K, int& s, int& nb) {
    if (K <= 12) {
        s = K;
        nb = 1;
        return;
    }
    for (int c : {9, 10, 8, 11, 12, 7}) {
        const int r = K % c;
        if (r == 0 || r >= 4) {
            s = c;
            nb = (K + c - 1) / c;
            return;
        }
    }
    s = 6;  // remainder merged into the last brick: 6..11 points
    nb = K / 6;
}

// polarisable-site index list and its host mirror
// the polarisable sites, and their grouping by molecule into preconditioner
// blocks of at most 4 sites (molecules in order of their first site)
static int set_polarisable(mbpol_context* ctx, Elec* E, const std::vector<double>& pol, const std::vector<int>& mol) {
    std::vector<int> idx;
    E->h_pol.assign(pol.size(), 0);
    std::vector<int> order;                      // molecules by first polarisable site
    std::map<int, std::vector<int>> by_mol;
    for (size_t i = 0; i < pol.size(); ++i)
        if (pol[i] != 0.0) {
            idx.push_back((int)i);
            E->h_pol[i] = 1;
            auto& v = by_mol[mol[i]];
            if (v.empty()) order.push_back(mol[i]);
            v.push_back((int)i);
        }
    E->npol = (int)idx.size();
    std::vector<int> start{0}, site;
    for (int m : order) {
        const std::vector<int>& v = by_mol[m];
        for (size_t a = 0; a < v.size(); a += 4) {
            for (size_t b = a; b < std::min(v.size(), a + 4); ++b) site.push_back(v[b]);
            start.push_back((int)site.size());
        }
    }
    E->nblk = (int)start.size() - 1;
    if (int e = up(ctx->stream, E->blk_start, start)) return e;
    if (int e = up(ctx->stream, E->blk_site, site)) return e;
    HIP_TRY(E->minv.ensure(sizeof(double) * block_ld() * block_ld() * std::max(E->nblk, 1)));
    return up(ctx->stream, E->polidx, idx);
}

int elec_set_box(mbpol_context* ctx) {
    Elec* E = ctx->elec;
    const mbpol_system& s = ctx->sys;
    ElecParams& P = E->P;
    P.box = d3{s.box[0], s.box[1], s.box[2]};
    P.inv_box = d3{s.box[0] > 0 ? 1.0 / s.box[0] : 0.0, s.box[1] > 0 ? 1.0 / s.box[1] : 0.0,
                   s.box[2] > 0 ? 1.0 / s.box[2] : 0.0};
    if (!P.pme) return MBPOL_OK;
    // cells of at least rc / reach, searched +-reach cells: half-cutoff cells
    // when the box holds five of them per axis, else cutoff cells, else none
    E->use_cells = 0;
    E->nc = make_int3(1, 1, 1);
    for (int reach = 2; reach >= 1 && !E->use_cells; --reach) {
        int nc[3];
        bool ok = true;
        for (int d = 0; d < 3; ++d) {
            nc[d] = (int)std::floor(s.box[d] / (P.cutoff / reach));
            ok = ok && nc[d] >= 2 * reach + 1;
        }
        if (ok) {
            E->use_cells = 1;
            E->reach = reach;
            E->nc = make_int3(nc[0], nc[1], nc[2]);
        }
    }
    const int ncell = E->nc.x * E->nc.y * E->nc.z;
    HIP_TRY(E->cell_count.ensure(sizeof(int) * (ncell + 1)));
    HIP_TRY(E->cell_start.ensure(sizeof(int) * (ncell + 1)));
    HIP_TRY(E->cell_fill.ensure(sizeof(int) * (ncell + 1)));
    // row capacities from the mean site density (grown on overflow)
    const double vol = s.box[0] * s.box[1] * s.box[2];
    const double mean = E->n / vol * (4.0 / 3.0) * M_PI * P.cutoff * P.cutoff * P.cutoff;
    auto cap = [](double v) { return std::max(32, ((int)std::ceil(v) + 31) / 32 * 32); };
    const double fp = E->n > 0 ? (double)E->npol / E->n : 0.0;
    E->capP = std::max(E->capP, cap(1.25 * mean * fp + 48));
    E->capQ = std::max(E->capQ, cap(1.25 * mean * (1.0 - fp) + 48));
    launch_influence(ctx->stream, P, E->fft.khp, E->bx.as<double>(), E->by.as<double>(), E->bz.as<double>(),
                     E->influence.as<double>());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MBPOL_OK;
}

// rows of the direct sum: every other site, or every other site of the group
static void set_direct_caps(Elec* E) {
    if (E->ngroups) {
        E->capP = std::max(E->cap_grp_pol, 1);
        E->capQ = std::max(E->cap_grp_other, 1);
    } else {
        E->capP = std::max(E->npol, 1);
        E->capQ = std::max(E->n - E->npol, 1);
    }
}

// independent groups of waters: site ranges (4 sites per water), one preconditioner block per water
static int set_groups(mbpol_context* ctx) {
    Elec* E = ctx->elec;
    const std::vector<int>& ws = ctx->sys.grp_wstart;
    E->ngroups = ws.empty() ? 0 : (int)ws.size() - 1;
    if (!E->ngroups) return MBPOL_OK;
    const int G = E->ngroups;
    std::vector<int> ss(E->n), se(E->n), bs(G + 1), ns(G);
    E->cap_grp_pol = E->cap_grp_other = 0;
    for (int g = 0; g < G; ++g) {
        int np = 0;
        for (int i = 4 * ws[g]; i < 4 * ws[g + 1]; ++i) {
            ss[i] = 4 * ws[g];
            se[i] = 4 * ws[g + 1];
            np += E->h_pol[i] ? 1 : 0;
        }
        ns[g] = 4 * (ws[g + 1] - ws[g]);
        bs[g] = ws[g];
        E->cap_grp_pol = std::max(E->cap_grp_pol, np);
        E->cap_grp_other = std::max(E->cap_grp_other, ns[g] - np);
    }
    bs[G] = ws[G];
    if (E->nblk != ws[G])
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "groups: every water needs its polarisable sites");
    if (int e = up(ctx->stream, E->g_site_s, ss)) return e;
    if (int e = up(ctx->stream, E->g_site_e, se)) return e;
    if (int e = up(ctx->stream, E->g_blk_start, bs)) return e;
    if (int e = up(ctx->stream, E->g_nsites, ns)) return e;
    HIP_TRY(E->g_blk_ee.ensure(sizeof(double) * std::max(E->nblk, 1)));
    HIP_TRY(E->g_esite.ensure(sizeof(double) * std::max(E->n, 1)));
    E->P.gs = E->g_site_s.as<int>();
    E->P.ge = E->g_site_e.as<int>();
    return MBPOL_OK;
}

int elec_create(mbpol_context* ctx) {
    const mbpol_system& s = ctx->sys;
    Elec* E = new Elec();
    ctx->elec = E;
    // device address of the no-allocation entry point's scratch array, once per device
    {
        double* sc = nullptr;
        if (int e = scratch_address(ctx->device, &sc)) return e;
    }
    E->n = s.ne;
    E->nw = s.el_redistribute ? s.ne / 4 : 0;
    // the FP64 mode runs the reference's SOR iteration, the faster modes conjugate gradients
    E->solver = ctx->precision == MBPOL_PRECISION_DOUBLE ? MBPOL_SOLVER_SOR : MBPOL_SOLVER_PCG;
    ElecParams& P = E->P;
    P.n = s.ne;
    P.pme = s.el_method == MBPOL_ELEC_PME;
    P.redistribute = s.el_redistribute;
    for (int k = 0; k < 5; ++k) P.thole[k] = s.el_thole[k];
    P.tcc_q = std::pow(P.thole[0], 0.25);
    P.cutoff = P.pme ? s.el_cutoff : 1e30;
    P.alpha = s.el_alpha;
    int grid[3] = {s.el_grid[0], s.el_grid[1], s.el_grid[2]};
    if (P.pme && (grid[0] == 0 || P.alpha == 0.0)) {
        // NonbondedForceImpl::calcPMEParameters as called at MBPolReferenceKernels.cpp:186-199
        const double tol = s.el_tol;
        P.alpha = (1.0 / s.el_cutoff) * std::sqrt(-std::log(2.0 * tol));
        for (int d = 0; d < 3; ++d)
            grid[d] = std::max((int)std::ceil(2 * P.alpha * s.box[d] / (3 * std::pow(tol, 0.2))), 6);
    }
    if (P.pme && (grid[0] < 5 || grid[1] < 5 || grid[2] < 5))
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "PME grid %d x %d x %d is smaller than the spline order 5",
                         grid[0], grid[1], grid[2]);
    P.grid = make_int3(grid[0], grid[1], grid[2]);
    const int n = E->n;
    if (int e = up(ctx->stream, E->q0, s.el_q)) return e;
    if (int e = up(ctx->stream, E->damp6, sixth_roots(s.el_damp))) return e;
    if (int e = up(ctx->stream, E->pol, s.el_pol)) return e;
    if (int e = up(ctx->stream, E->mol, s.el_mol)) return e;
    if (int e = up(ctx->stream, E->type, s.el_type)) return e;
    if (int e = set_polarisable(ctx, E, s.el_pol, s.el_mol)) return e;
    HIP_TRY(E->q.ensure(sizeof(double) * n));
    HIP_TRY(E->dq.ensure(sizeof(double) * 27 * std::max(E->nw, 1)));
    for (DevBuf* b : {&E->efix, &E->efix_pol, &E->mu_a, &E->mu_b, &E->mu_prev, &E->mu_user, &E->fsite})
        HIP_TRY(b->ensure(sizeof(double) * 3 * std::max(n, 1)));
    // sites the iteration never visits keep zero dipoles in both buffers
    launch_fill_zero(ctx->stream, E->mu_a.p, sizeof(double) * 3 * std::max(n, 1));
    launch_fill_zero(ctx->stream, E->mu_b.p, sizeof(double) * 3 * std::max(n, 1));
    launch_fill_zero(ctx->stream, E->mu_prev.p, sizeof(double) * 3 * std::max(n, 1));
    E->mu_cur = E->mu_a.as<double>();
    HIP_TRY(E->V.ensure(sizeof(double) * std::max(n, 1)));
    HIP_TRY(E->phi_fix.ensure(sizeof(double) * 10 * std::max(n, 1)));
    HIP_TRY(E->phi_ind.ensure(sizeof(double) * 10 * std::max(n, 1)));
    HIP_TRY(E->partial.ensure(sizeof(double) * (iterate_blocks(E->npol) + 1)));
    HIP_TRY(E->eps.ensure(sizeof(double) * 4));
    HIP_TRY(E->maxcnt.ensure(sizeof(int) * 4));
    HIP_TRY(E->eblk.ensure(sizeof(double) * (std::max(energy_blocks(n), sor_blocks(n)) + 1)));
    HIP_TRY(E->cell_of.ensure(sizeof(int) * std::max(n, 1)));
    HIP_TRY(E->members.ensure(sizeof(int) * std::max(n, 1)));
    HIP_TRY(E->cpos.ensure(sizeof(double4) * std::max(n, 1)));
    HIP_TRY(E->unsorted.ensure(sizeof(int) * std::max(n, 1)));
    HIP_TRY(E->cntP.ensure(sizeof(int) * std::max(n, 1)));
    HIP_TRY(E->cntQ.ensure(sizeof(int) * std::max(n, 1)));
    HIP_TRY(E->cell_count.ensure(sizeof(int) * 2));
    HIP_TRY(E->cell_start.ensure(sizeof(int) * 2));
    HIP_TRY(E->cell_fill.ensure(sizeof(int) * 2));
    HIP_TRY(hipHostMalloc((void**)&E->h_eps, 128, hipHostMallocDefault));
    HIP_TRY(E->ctl.ensure(sizeof(SolveCtl)));
    if (int e = set_groups(ctx)) return e;
    if (!P.pme) set_direct_caps(E);  // direct sum: rows hold every other site (of the group)
    if (P.pme) {
        const size_t nreal = (size_t)grid[0] * grid[1] * grid[2];
        bool lds = true;
        for (int d = 0; d < 3; ++d) lds = lds && fft_factor(grid[d], E->fft.ax[d]);
        const char* force = std::getenv("MBPOL_HIP_PME_FFT");  // "hipfft" selects the library transforms
        if (force && std::string(force) == "hipfft") lds = false;
        E->lds_fft = lds;
        // z and y passes fused per x plane: faster alone, but a plane workgroup takes a
        // whole CU and stalls beside the short-range kernels on the second stream,
        // so the five single-axis passes are the default (MBPOL_HIP_PME_PLANES=1: planes)
        const char* planes = std::getenv("MBPOL_HIP_PME_PLANES");
        E->plane_fft = lds && planes && std::string(planes) == "1" && fft_plane_ok(E->fft);
        const char* fuse = std::getenv("MBPOL_HIP_SPREAD_FUSED");
        E->fuse_spread = !(fuse && std::string(fuse) == "0");
        // half-spectrum rows: padded to 128 bytes for the LDS transforms, packed for hipFFT
        const int kh = grid[2] / 2 + 1;
        E->fft.khp = lds ? (kh + 7) / 8 * 8 : kh;
        const size_t nhalf = (size_t)grid[0] * grid[1] * E->fft.khp;
        HIP_TRY(E->grid.ensure(sizeof(double) * nreal));
        HIP_TRY(E->spec.ensure(sizeof(double2) * nhalf));
        HIP_TRY(E->influence.ensure(sizeof(double) * nhalf));
        if (int e = up(ctx->stream, E->bx, bspline_moduli(grid[0]))) return e;
        if (int e = up(ctx->stream, E->by, bspline_moduli(grid[1]))) return e;
        if (int e = up(ctx->stream, E->bz, bspline_moduli(grid[2]))) return e;
        for (int d = 0; d < 3; ++d) brick_axis(grid[d], E->B.s[d], E->B.nb[d]);
        const int nbins = E->B.nb[0] * E->B.nb[1] * E->B.nb[2];
        HIP_TRY(E->binq.of.ensure(sizeof(int) * std::max(n, 1)));
        HIP_TRY(E->binq.members.ensure(sizeof(int) * std::max(n, 1)));
        HIP_TRY(E->binq.count.ensure(sizeof(int) * (nbins + 1)));
        HIP_TRY(E->binq.start.ensure(sizeof(int) * (nbins + 1)));
        HIP_TRY(E->binq.fill.ensure(sizeof(int) * (nbins + 1)));
        HIP_TRY(E->g0tab.ensure(sizeof(int) * 3 * std::max(n, 1)));
        HIP_TRY(E->sptab.ensure(sizeof(double) * 30 * std::max(n, 1)));
        for (DevBuf* b : {&E->rowcount, &E->rowstart, &E->rowfill}) HIP_TRY(b->ensure(sizeof(int) * (nreal + 1)));
        if (lds) {
            for (int d = 0; d < 3; ++d) {
                const int m = grid[d];
                std::vector<double2> tw(m);
                for (int t = 0; t < m; ++t) {
                    const long double a = 2.0L * 3.141592653589793238462643383279502884L * t / m;
                    tw[t] = make_double2((double)cosl(a), (double)-sinl(a));
                }
                if (int e = up(ctx->stream, E->tw[d], tw)) return e;
                E->fft.ax[d].tw = E->tw[d].as<double2>();
            }
        } else {
            FFT_TRY(hipfftPlan3d(&E->plan_f, grid[0], grid[1], grid[2], HIPFFT_D2Z));
            FFT_TRY(hipfftPlan3d(&E->plan_b, grid[0], grid[1], grid[2], HIPFFT_Z2D));
            E->plans = true;
            FFT_TRY(hipfftSetStream(E->plan_f, ctx->stream));
            FFT_TRY(hipfftSetStream(E->plan_b, ctx->stream));
        }
    }
    return elec_set_box(ctx);
}

void elec_destroy(mbpol_context* ctx) {
    Elec* E = ctx->elec;
    if (!E) return;
    if (E->plans) {
        hipfftDestroy(E->plan_f);
        hipfftDestroy(E->plan_b);
    }
    if (E->h_eps) (void)hipHostFree(E->h_eps);
    delete E;
    ctx->elec = nullptr;
}

static ElecRows rows(const Elec* E) {
    return ElecRows{E->capP, E->capQ, E->nbP.as<int>(), E->cntP.as<int>(), E->nbQ.as<int>(), E->cntQ.as<int>(),
                    E->pf.as<double2>()};
}

// neighbour rows of every site (PME: within the cutoff; direct sum: every
// other site), the induced-field prefactors, and the PME bins
// Autotuning of the row build: the lane count per site (8, 16, 32) is timed on
// this context's own first configuration, each candidate twice (the second
// launch counts), and the fastest is kept.  Every candidate writes the same
// rows.  MBPOL_HIP_SITE_LANES=8|16|32 fixes the choice instead.
static int tune_site_lists(mbpol_context* ctx) {
    Elec* E = ctx->elec;
    if (const char* env = std::getenv("MBPOL_HIP_SITE_LANES")) {
        const int v = std::atoi(env);
        E->site_lanes = (v == 8 || v == 32) ? v : 16;
        return MBPOL_OK;
    }
    if (!E->use_cells) {  // direct sum: rows of every site, no search to tune
        E->site_lanes = 16;
        return MBPOL_OK;
    }
    hipStream_t st = ctx->stream;
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    float best = 1e30f;
    int choice = 16;
    for (int lanes : {8, 16, 32}) {
        float ms = 0.f;
        for (int rep = 0; rep < 2; ++rep) {
            launch_fill_zero(st, E->maxcnt.p, sizeof(int) * 4);
            HIP_TRY(hipEventRecord(a, st));
            launch_site_lists(st, lanes, E->n, ctx->pos.as<double>(), E->pol.as<double>(), E->P, E->nc, E->reach,
                              E->use_cells, E->cell_of.as<int>(), E->cell_start.as<int>(), E->members.as<int>(),
                              E->cpos.as<double4>(), E->capP, E->capQ, E->nbP.as<int>(), E->cntP.as<int>(),
                              E->nbQ.as<int>(), E->cntQ.as<int>(), E->maxcnt.as<int>());
            HIP_TRY(hipEventRecord(b, st));
            HIP_TRY(hipEventSynchronize(b));
            HIP_TRY(hipEventElapsedTime(&ms, a, b));
        }
        if (ms < best) {
            best = ms;
            choice = lanes;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    E->site_lanes = choice;
    return MBPOL_OK;
}

int elec_tuning(const mbpol_context* ctx, const char* name, int* value) {
    const Elec* E = ctx->elec;
    if (E && std::strcmp(name, "site_lanes") == 0)
        *value = E->site_lanes;
    else if (E && std::strcmp(name, "row_capacity_pol") == 0)  // entries per polarisable-neighbour row
        *value = E->capP;
    else if (E && std::strcmp(name, "row_capacity_other") == 0)  // entries per row of the other neighbours
        *value = E->capQ;
    else
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "no tuning '%s'", name);
    return MBPOL_OK;
}

static int build_rows(mbpol_context* ctx) {
    Elec* E = ctx->elec;
    hipStream_t st = ctx->stream;
    const int n = E->n;
    const double* pos = ctx->pos.as<double>();
    const int ncell = E->nc.x * E->nc.y * E->nc.z;
    if (E->use_cells) {
        launch_site_cells(st, n, pos, E->P, E->nc, E->cell_of.as<int>(), E->cell_count.as<int>());
        if (int e = scan_i32(ctx, E->cell_count.as<int>(), E->cell_start.as<int>(), ncell + 1)) return e;
        launch_bucket_fill(st, n, E->cell_of.as<int>(), E->cell_start.as<int>(), E->cell_fill.as<int>(),
                           E->unsorted.as<int>());
        launch_rank_sort(st, n, E->cell_of.as<int>(), E->cell_start.as<int>(), E->unsorted.as<int>(),
                         E->members.as<int>());
        launch_site_pack(st, n, E->members.as<int>(), pos, E->pol.as<double>(), E->cpos.as<double4>());
    }
    int* h = reinterpret_cast<int*>(E->h_eps) + 8;
    E->rows_pending = false;
    if (!E->rows_sync && E->site_lanes > 0) {
        // The rows' maxima reach the host with the energies; elec_rows_check
        // compares them with the capacities after the evaluation and has it
        // redone when a row overflowed (the kernels clamp to the capacity, so an
        // overflowing row loses neighbours but nothing is written out of bounds).
        const size_t rp = (size_t)n * E->capP, rq = (size_t)n * E->capQ;
        HIP_TRY(E->nbP.ensure(sizeof(int) * std::max<size_t>(rp, 1)));
        HIP_TRY(E->pf.ensure(sizeof(double2) * std::max<size_t>(rp, 1)));
        HIP_TRY(E->nbQ.ensure(sizeof(int) * std::max<size_t>(rq, 1)));
        launch_site_lists(st, E->site_lanes, n, pos, E->pol.as<double>(), E->P, E->nc, E->reach, E->use_cells,
                          E->cell_of.as<int>(), E->cell_start.as<int>(), E->members.as<int>(), E->cpos.as<double4>(),
                          E->capP, E->capQ, E->nbP.as<int>(), E->cntP.as<int>(), E->nbQ.as<int>(), E->cntQ.as<int>(),
                          E->maxcnt.as<int>());
        HIP_TRY(hipMemcpyAsync(h, E->maxcnt.p, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
        E->rows_pending = true;
    } else for (int attempt = 0; attempt < 2; ++attempt) {
        const size_t rp = (size_t)n * E->capP, rq = (size_t)n * E->capQ;
        HIP_TRY(E->nbP.ensure(sizeof(int) * std::max<size_t>(rp, 1)));
        HIP_TRY(E->pf.ensure(sizeof(double2) * std::max<size_t>(rp, 1)));
        HIP_TRY(E->nbQ.ensure(sizeof(int) * std::max<size_t>(rq, 1)));
        launch_fill_zero(st, E->maxcnt.p, sizeof(int) * 4);
        launch_site_lists(st, E->site_lanes > 0 ? E->site_lanes : 16, n, pos, E->pol.as<double>(), E->P, E->nc, E->reach,
                          E->use_cells, E->cell_of.as<int>(), E->cell_start.as<int>(), E->members.as<int>(),
                          E->cpos.as<double4>(), E->capP, E->capQ, E->nbP.as<int>(), E->cntP.as<int>(),
                          E->nbQ.as<int>(), E->cntQ.as<int>(), E->maxcnt.as<int>());
        HIP_TRY(hipMemcpyAsync(h, E->maxcnt.p, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (h[0] <= E->capP && h[1] <= E->capQ) break;
        if (attempt == 1) return set_error(MBPOL_ERR_HIP, "electrostatics neighbour rows overflowed twice");
        E->capP = std::max(E->capP, (h[0] + h[0] / 4 + 63) / 32 * 32);
        E->capQ = std::max(E->capQ, (h[1] + h[1] / 4 + 63) / 32 * 32);
    }
    if (E->site_lanes == 0)
        if (int e = tune_site_lists(ctx)) return e;
    if (!E->rows_pending) {
        E->rows_sync = false;
        ctx->stats.elec_pairs = h[2] / 2;
        ctx->stats.pol_pairs = h[3];
    }
    // prefactors of the P rows and the real-space part of the fixed field (the
    // reciprocal part is added to it after the charge grid)
    launch_prefactors_field(st, E->npol, E->polidx.as<int>(), pos, site_params(E), E->P, rows(E), E->pf.as<double2>(),
                            E->efix.as<double>());
    if (E->P.pme) {
        const int nbins = E->B.nb[0] * E->B.nb[1] * E->B.nb[2];
        launch_site_spline_table(st, n, pos, E->P, E->g0tab.as<int>(), E->sptab.as<double>());
        // dipole-spreading table
        const size_t ngrid = (size_t)E->P.grid.x * E->P.grid.y * E->P.grid.z;
        const size_t nent = (size_t)E->npol * 125;
        HIP_TRY(E->esite.ensure(sizeof(int) * std::max<size_t>(nent, 1)));
        HIP_TRY(E->ecoef.ensure(sizeof(double) * 3 * std::max<size_t>(nent, 1)));
        launch_csr_count(st, E->npol, E->polidx.as<int>(), E->g0tab.as<int>(), E->P, E->rowcount.as<int>());
        if (int e = scan_i32(ctx, E->rowcount.as<int>(), E->rowstart.as<int>(), (int)ngrid + 1)) return e;
        launch_csr_fill(st, E->npol, E->polidx.as<int>(), E->g0tab.as<int>(), E->sptab.as<double>(), E->P,
                        E->rowstart.as<int>(), E->rowfill.as<int>(), E->esite.as<int>(), E->ecoef.as<double>());
        launch_csr_sort(st, ngrid, E->rowstart.as<int>(), E->esite.as<int>(), E->ecoef.as<double>());
        // charged sites by brick, for the charge spreading
        Elec::Bins& b = E->binq;
        launch_site_bins(st, n, pos, E->q.as<double>(), E->P, E->B, b.of.as<int>(), b.count.as<int>());
        if (int e = scan_i32(ctx, b.count.as<int>(), b.start.as<int>(), nbins + 1)) return e;
        launch_bucket_fill(st, n, b.of.as<int>(), b.start.as<int>(), b.fill.as<int>(), E->unsorted.as<int>());
        launch_rank_sort(st, n, b.of.as<int>(), b.start.as<int>(), E->unsorted.as<int>(), b.members.as<int>());
    }
    return MBPOL_OK;
}

// reciprocal-space potential on the grid of the charges (mode 0) or of the dipoles mu (mode 1)
static int pme_grid(mbpol_context* ctx, int mode, const double* mu, const int* stop = nullptr) {
    Elec* E = ctx->elec;
    hipStream_t st = ctx->stream;
    const ElecParams& P = E->P;
    const size_t nhalf = (size_t)P.grid.x * P.grid.y * (P.grid.z / 2 + 1);
    // dipoles through the single-axis LDS passes: spreading and z pass in one kernel
    const bool fused = mode == 1 && E->lds_fft && !E->plane_fft && E->fuse_spread;
    if (mode == 0) {
        launch_spread_bricks(st, mode, E->q.as<double>(), mu, P, E->B, E->binq.start.as<int>(),
                             E->binq.members.as<int>(), E->g0tab.as<int>(), E->sptab.as<double>(), E->grid.as<double>());
    } else if (fused) {
        const KEvents ev = timer_kernel(ctx, T_SPREAD_FFT_Z);
        launch_spread_fft_z_r2c(st, E->fft, E->rowstart.as<int>(), E->esite.as<int>(), E->ecoef.as<double>(), mu,
                                E->spec.as<double2>(), ev, stop);
        timer_kernel_done(ctx, T_SPREAD_FFT_Z, ev);
    } else {
        const KEvents ev = timer_kernel(ctx, T_PME_SPREAD);
        launch_spread_csr(st, (size_t)P.grid.x * P.grid.y * P.grid.z, E->rowstart.as<int>(), E->esite.as<int>(),
                          E->ecoef.as<double>(), mu, E->grid.as<double>(), ev, stop);
        timer_kernel_done(ctx, T_PME_SPREAD, ev);
    }
    if (E->lds_fft) {
        double2* H = E->spec.as<double2>();
        const bool planes = E->plane_fft;
        KEvents ev;
        if (planes) {
            ev = timer_kernel(ctx, T_FFT_PLANE_FWD);
            launch_fft_plane_fwd(st, E->fft, E->grid.as<double>(), H, ev, stop);
            timer_kernel_done(ctx, T_FFT_PLANE_FWD, ev);
        } else {
            if (!fused) {
                ev = timer_kernel(ctx, T_FFT_Z_R2C);
                launch_fft_z_r2c(st, E->fft, E->grid.as<double>(), H, ev, stop);
                timer_kernel_done(ctx, T_FFT_Z_R2C, ev);
            }
            ev = timer_kernel(ctx, T_FFT_Y);
            launch_fft_y(st, E->fft, H, 0, ev, stop);
            timer_kernel_done(ctx, T_FFT_Y, ev);
        }
        ev = timer_kernel(ctx, T_FFT_X_CONV);
        launch_fft_x_conv(st, E->fft, H, E->influence.as<double>(), ev, stop);
        timer_kernel_done(ctx, T_FFT_X_CONV, ev);
        if (planes) {
            ev = timer_kernel(ctx, T_FFT_PLANE_INV);
            launch_fft_plane_inv(st, E->fft, H, E->grid.as<double>(), ev, stop);
            timer_kernel_done(ctx, T_FFT_PLANE_INV, ev);
        } else {
            ev = timer_kernel(ctx, T_FFT_Y);
            launch_fft_y(st, E->fft, H, 1, ev, stop);
            timer_kernel_done(ctx, T_FFT_Y, ev);
            ev = timer_kernel(ctx, T_FFT_Z_C2R);
            launch_fft_z_c2r(st, E->fft, H, E->grid.as<double>(), ev, stop);
            timer_kernel_done(ctx, T_FFT_Z_C2R, ev);
        }
    } else {
        FFT_TRY(hipfftExecD2Z(E->plan_f, E->grid.as<double>(), reinterpret_cast<hipfftDoubleComplex*>(E->spec.p)));
        launch_convolve(st, nhalf, E->influence.as<double>(), E->spec.as<double2>());
        FFT_TRY(hipfftExecZ2D(E->plan_b, reinterpret_cast<hipfftDoubleComplex*>(E->spec.p), E->grid.as<double>()));
    }
    return MBPOL_OK;
}

// The dipole solves enqueue their steps in batches without waiting for the
// convergence test: each step's last kernel applies the stopping rules on the
// device (SolveCtl), and every kernel of a later step returns at once when the
// solve has stopped.  The host reads the state once per batch.  The first batch
// covers the previous solve's length (a solve as long as the last one ends with
// one read and no step that returns at once); later batches take 1, 2, 4, ...
// steps (a step that only returns costs ~8 launches, a read ~35 us).  With
// timing on, or with hipFFT transforms (which cannot skip), one step per batch.
static int solve_batch(const Elec* E, const mbpol_context* ctx, int batch) {
    if (ctx->timer.enabled || !E->lds_fft) return 1;
    if (batch == 0) return std::max(1, E->last_iters);
    return 1 << std::min(batch - 1, 4);
}

static int read_ctl(mbpol_context* ctx, SolveCtl& out) {
    Elec* E = ctx->elec;
    SolveCtl* h = reinterpret_cast<SolveCtl*>(reinterpret_cast<char*>(E->h_eps) + 64);
    HIP_TRY(hipMemcpyAsync(h, E->ctl.p, sizeof(SolveCtl), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    out = *h;
    return MBPOL_OK;
}

// SOR iteration (convergeInduceDipoles, :557-616) from the guess in buf[0];
// step t reads buf[t % 2] and writes buf[(t + 1) % 2].  On return mu_in holds
// the dipoles after the last step and mu_out those before it.
static int solve_sor(mbpol_context* ctx, double* buf0, double* buf1, double self_term, double*& mu_in,
                     double*& mu_out, SolveCtl& c) {
    Elec* E = ctx->elec;
    const mbpol_system& s = ctx->sys;
    hipStream_t st = ctx->stream;
    const ElecParams& P = E->P;
    const ElecRows R = rows(E);
    const double* pos = ctx->pos.as<double>();
    SolveCtl* ctl = E->ctl.as<SolveCtl>();
    const int* stop = &ctl->stop;
    double* buf[2] = {buf0, buf1};
    int t = 0, batch = 0;
    c = SolveCtl{};
    while (t <= s.el_max_iter) {
        const int nb = std::min(solve_batch(E, ctx, batch++), s.el_max_iter + 1 - t);
        for (int b = 0; b < nb; ++b, ++t) {
            timer_start(ctx, T_ELEC_ITER);
            if (P.pme)
                if (int e = pme_grid(ctx, 1, buf[t % 2], stop)) return e;
            const KEvents ev = timer_kernel(ctx, T_ELEC_ITERATE);
            launch_iterate(st, E->npol, E->polidx.as<int>(), pos, P, R, buf[t % 2], buf[(t + 1) % 2],
                           E->efix_pol.as<double>(), E->pol.as<double>(), E->grid.as<double>(), self_term,
                           mbpol_params::el_polar_sor, E->partial.as<double>(), ctl, ev);
            timer_kernel_done(ctx, T_ELEC_ITERATE, ev);
            launch_sor_finish(st, iterate_blocks(E->npol), E->n, E->partial.as<double>(), t, s.el_eps, s.el_max_iter,
                              ctl);
            timer_stop(ctx, T_ELEC_ITER);
        }
        if (int e = read_ctl(ctx, c)) return e;
        if (c.stop) break;
    }
    const int it = c.iterations;
    mu_in = buf[it % 2];
    mu_out = buf[(it + 1) % 2];
    return MBPOL_OK;
}

// Preconditioned conjugate gradients on the symmetric form of the dipole
// equations (kernels_elec.hip, k_cg_*), started from the guess in x.  Each
// step costs one induced-field evaluation, as an SOR step does; the stopping
// measure is the reference's epsilon of the current dipoles, tested after every
// step against the same target.  Iterations count field evaluations (the
// initial residual included).
// fixed_steps > 0: exactly that many field evaluations (or max_iter + 1), with
// no host read; the stopping rule on the device makes the steps after the stop
// return at once, and the state is copied to the host for reading after the
// work has run (elec_graph_finish).  This form can be captured in a HIP graph.
static int solve_pcg(mbpol_context* ctx, double* x, double self_term, SolveCtl& c, int fixed_steps = 0) {
    Elec* E = ctx->elec;
    const mbpol_system& s = ctx->sys;
    hipStream_t st = ctx->stream;
    const ElecParams& P = E->P;
    const ElecRows R = rows(E);
    const double* pos = ctx->pos.as<double>();
    const int n = E->n, npol = E->npol;
    c = SolveCtl{};
    if (npol == 0) {  // nothing to polarise
        c.converged = 1;
        return MBPOL_OK;
    }
    const size_t vbytes = sizeof(double) * 3 * std::max(n, 1);
    for (DevBuf* b : {&E->cg_r, &E->cg_z, &E->cg_p, &E->cg_q}) HIP_TRY(b->ensure(vbytes));
    HIP_TRY(E->cg_partial.ensure(sizeof(double) * 2 * (precond_wgs(E->nblk) + 1)));
    HIP_TRY(E->cg_rz.ensure(sizeof(double) * (s.el_max_iter + 2)));
    HIP_TRY(E->partial.ensure(sizeof(double) * 2 * (iterate_blocks(npol) + 1)));
    double *r = E->cg_r.as<double>(), *z = E->cg_z.as<double>(), *p = E->cg_p.as<double>(), *q = E->cg_q.as<double>();
    double* rz = E->cg_rz.as<double>();
    const double* pol = E->pol.as<double>();
    const int* polidx = E->polidx.as<int>();
    SolveCtl* ctl = E->ctl.as<SolveCtl>();
    const int* stop = &ctl->stop;
    const PrecondBlocks B{E->nblk, E->blk_start.as<int>(), E->blk_site.as<int>()};
    double* minv = E->minv.as<double>();
    // independent groups: convergence of the worst group
    const GroupConv gc{E->ngroups, E->g_blk_start.as<int>(), E->g_nsites.as<int>(), E->g_blk_ee.as<double>()};
    double* blk_ee = E->ngroups ? E->g_blk_ee.as<double>() : nullptr;
    auto field = [&](const double* v, int mode, double* out) -> int {
        if (P.pme)
            if (int e = pme_grid(ctx, 1, v, stop)) return e;
        const KEvents ev = timer_kernel(ctx, T_ELEC_ITERATE);
        launch_cg_matvec(st, npol, polidx, pos, P, R, v, E->grid.as<double>(), self_term, pol, E->efix.as<double>(),
                         mode, out, E->partial.as<double>(), ctl, ev);
        timer_kernel_done(ctx, T_ELEC_ITERATE, ev);
        return MBPOL_OK;
    };
    launch_block_factor(st, B, pos, site_params(E), P, minv);
    // field evaluation f (0-based): f = 0 forms the initial residual and direction 0,
    // f = k + 1 takes step k and forms direction k + 1
    int f = 0, batch = 0;
    while (f <= s.el_max_iter) {
        const int nb = fixed_steps > 0 ? std::min(fixed_steps, s.el_max_iter + 1)
                                       : std::min(solve_batch(E, ctx, batch++), s.el_max_iter + 1 - f);
        for (int b = 0; b < nb; ++b, ++f) {
            timer_start(ctx, T_ELEC_ITER);
            if (f == 0) {
                if (int e = field(x, 0, r)) return e;
                launch_cg_precond(st, B, minv, pol, r, z, E->cg_partial.as<double>(), blk_ee, ctl);
                launch_cg_direction(st, npol, polidx, precond_wgs(E->nblk), E->cg_partial.as<double>(), rz, 0, z, p, n,
                                    s.el_eps, s.el_max_iter, gc, ctl);
            } else {
                const int k = f - 1;
                if (int e = field(p, 1, q)) return e;
                launch_cg_update(st, B, minv, pol, npol, E->partial.as<double>(), rz, k, x, r, p, q, z,
                                 E->cg_partial.as<double>(), blk_ee, ctl);
                launch_cg_direction(st, npol, polidx, precond_wgs(E->nblk), E->cg_partial.as<double>(), rz, k + 1, z, p,
                                    n, s.el_eps, s.el_max_iter, gc, ctl);
            }
            timer_stop(ctx, T_ELEC_ITER);
        }
        if (fixed_steps > 0) {
            HIP_TRY(hipMemcpyAsync(reinterpret_cast<char*>(E->h_eps) + 64, E->ctl.p, sizeof(SolveCtl),
                                   hipMemcpyDeviceToHost, st));
            return MBPOL_OK;
        }
        if (int e = read_ctl(ctx, c)) return e;
        if (c.stop) break;
    }
    return MBPOL_OK;
}

bool elec_graph_ok(const mbpol_context* ctx) {
    const Elec* E = ctx->elec;
    return E->solver == MBPOL_SOLVER_PCG && !E->rows_sync && E->site_lanes > 0 && (!E->P.pme || E->lds_fft) &&
           E->guess_mode != MBPOL_GUESS_LEARNED && E->history == 0;
}

int elec_graph_steps(const mbpol_context* ctx, int extra) {
    const Elec* E = ctx->elec;
    return std::min(ctx->sys.el_max_iter + 1, std::max(E->last_iters, 1) + extra);
}

bool elec_tune_ok(const mbpol_context* ctx) {
    const Elec* E = ctx->elec;
    return E->history == 0 && E->guess_mode != MBPOL_GUESS_LEARNED;
}

int elec_tune_hold(mbpol_context* ctx, bool hold) {
    Elec* E = ctx->elec;
    hipStream_t st = ctx->stream;
    const size_t bytes = sizeof(double) * 3 * std::max(E->n, 1);
    if (hold) {
        HIP_TRY(E->mu_held.ensure(bytes));
        HIP_TRY(hipMemcpyAsync(E->mu_held.p, E->mu_prev.p, bytes, hipMemcpyDeviceToDevice, st));
        E->held_have_previous = E->have_previous;
        E->held_last_iters = E->last_iters;
        E->held_guess = E->guess_mode;
        E->guess_mode = MBPOL_GUESS_ALPHA_E;  // every trial solve starts alike
    } else {
        HIP_TRY(hipMemcpyAsync(E->mu_prev.p, E->mu_held.p, bytes, hipMemcpyDeviceToDevice, st));
        E->have_previous = E->held_have_previous;
        E->last_iters = E->held_last_iters;
        E->guess_mode = E->held_guess;
    }
    return MBPOL_OK;
}

void elec_graph_key(const mbpol_context* ctx, long long key[4]) {
    const Elec* E = ctx->elec;
    key[0] = E->have_previous;
    key[1] = E->guess_mode;
    key[2] = E->capP;
    key[3] = E->capQ;
}

int elec_graph_finish(mbpol_context* ctx, bool& more) {
    Elec* E = ctx->elec;
    more = false;
    const SolveCtl c = *reinterpret_cast<const SolveCtl*>(reinterpret_cast<const char*>(E->h_eps) + 64);
    E->rows_pending = true;  // the row maxima were copied to the host with the rest
    E->mu_cur = E->mu_a.as<double>();  // where the conjugate-gradient solve leaves the dipoles
    if (!c.stop) {  // the fixed steps did not reach the stopping rule
        more = true;
        return MBPOL_OK;
    }
    E->last_iters = c.iterations;
    ctx->stats.iterations = c.iterations;
    ctx->stats.epsilon = c.eps;
    ctx->stats.converged = c.converged ? 1 : 0;
    if (!c.converged)
        return set_error(MBPOL_ERR_NOT_CONVERGED, "Induced dipoles did not converge:  iterations=%d eps=%g",
                         c.iterations, c.eps);
    E->have_previous = true;
    return MBPOL_OK;
}

void elec_keep_history(mbpol_context* ctx, bool on) {
    ++ctx->graph_gen;
    Elec* E = ctx->elec;
    if (!E) return;
    if (on) {
        if (E->history++ == 0) {
            const size_t bytes = sizeof(double) * 3 * std::max(E->n, 1);
            if (E->efix_prev.ensure(bytes) || E->mu_prev2.ensure(bytes) || E->efix_prev2.ensure(bytes)) {
                --E->history;
                return;
            }
            // the history starts with the next evaluation
            E->have_previous = false;
            E->have_previous2 = false;
        }
    } else if (E->history > 0) {
        --E->history;
    }
}

int elec_rows_check(mbpol_context* ctx, bool& redo) {
    Elec* E = ctx->elec;
    redo = false;
    if (!E || !E->rows_pending) return MBPOL_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    E->rows_pending = false;
    const int* h = reinterpret_cast<const int*>(E->h_eps) + 8;
    if (h[0] > E->capP || h[1] > E->capQ) {
        E->capP = std::max(E->capP, (h[0] + h[0] / 4 + 63) / 32 * 32);
        E->capQ = std::max(E->capQ, (h[1] + h[1] / 4 + 63) / 32 * 32);
        E->rows_sync = true;
        redo = true;
        return MBPOL_OK;
    }
    ctx->stats.elec_pairs = h[2] / 2;
    ctx->stats.pol_pairs = h[3];
    return MBPOL_OK;
}

int elec_view(mbpol_context* ctx, ElecView& v) {
    Elec* E = ctx->elec;
    if (!E) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "no electrostatics in this system");
    const mbpol_system& s = ctx->sys;
    v.n = E->n;
    v.nw = E->n / 4;
    v.water_layout = s.el_redistribute && E->n % 4 == 0;
    for (int i = 0; v.water_layout && i < E->n; ++i)
        v.water_layout = s.el_type[i] == (i % 4 == 0 ? 0 : (i % 4 == 3 ? 2 : 1));
    v.efix_pol = E->efix_pol.as<double>();
    v.mu_prev = E->mu_prev.as<double>();
    v.have_previous = E->have_previous;
    if (E->history) {
        v.efix_prev = E->efix_prev.as<double>();
        v.mu_prev2 = E->mu_prev2.as<double>();
        v.efix_prev2 = E->efix_prev2.as<double>();
        v.have_previous2 = E->have_previous2;
    } else {
        v.have_previous = false;  // no alpha E_fixed of the last evaluation without the history
    }
    v.mu = E->mu_cur;
    v.type = E->type.as<int>();
    v.cell_of = E->cell_of.as<int>();
    v.cell_start = E->cell_start.as<int>();
    v.members = E->members.as<int>();
    v.use_cells = E->use_cells;
    v.nc = E->nc;
    v.periodic = E->P.pme;
    v.box = E->P.box;
    v.inv_box = E->P.inv_box;
    // cells reach the feature cutoff (0.6 nm) at reach ceil(0.6 / cell) without wrapping onto themselves
    v.reach = 0;
    if (E->use_cells) {
        const double cell = std::min({E->P.box.x / E->nc.x, E->P.box.y / E->nc.y, E->P.box.z / E->nc.z});
        const int r = (int)std::ceil(0.6 / cell);
        if (2 * r + 1 <= std::min({E->nc.x, E->nc.y, E->nc.z})) v.reach = r;
        else v.use_cells = 0;
    }
    return MBPOL_OK;
}

int elec_evaluate(mbpol_context* ctx, u64* energy, const std::function<int()>& after_setup, int fixed_steps) {
    Elec* E = ctx->elec;
    const mbpol_system& s = ctx->sys;
    hipStream_t st = ctx->stream;
    const int n = E->n;
    const ElecParams& P = E->P;
    const double* pos = ctx->pos.as<double>();
    const SiteParams S = site_params(E);
    // charges (setup, :857-897)
    timer_start(ctx, T_ELEC_SETUP);
    // every counter and accumulator of this evaluation cleared by one launch: cell and
    // row counters, the fixed field, the PME spreading-table and brick counters, the
    // solver state, and the energy phase's site forces and potentials
    {
        FillList L(st);
        const int ncell = E->nc.x * E->nc.y * E->nc.z;
        if (E->use_cells) {
            L.add(E->cell_count.p, sizeof(int) * (ncell + 1));
            L.add(E->cell_fill.p, sizeof(int) * (ncell + 1));
        }
        L.add(E->maxcnt.p, sizeof(int) * 4);
        L.add(E->efix.p, sizeof(double) * 3 * n);
        if (P.pme) {
            const size_t ngrid = (size_t)P.grid.x * P.grid.y * P.grid.z;
            const int nbins = E->B.nb[0] * E->B.nb[1] * E->B.nb[2];
            L.add(E->rowcount.p, sizeof(int) * (ngrid + 1));
            L.add(E->rowfill.p, sizeof(int) * (ngrid + 1));
            L.add(E->binq.count.p, sizeof(int) * (nbins + 1));
            L.add(E->binq.fill.p, sizeof(int) * (nbins + 1));
        }
        L.add(E->ctl.p, sizeof(SolveCtl));
        L.add(E->fsite.p, sizeof(double) * 3 * n);
        L.add(E->V.p, sizeof(double) * n);
        L.launch();
    }
    launch_site_charges_copy(st, n, E->q0.as<double>(), E->q.as<double>());
    // sites of water w are 4w..4w+3 (the reference's ii += 4 loop, :881)
    if (P.redistribute) {
        // with the 1-body term of the same waters when the evaluation runs it here
        const bool one = ctx->onebody_with_charges;
        const int G = ctx->ngroups;
        launch_water_terms(st, E->nw, one, true, ctx->ob_idx.as<int>(), pos, s.ob_method == MBPOL_ONEBODY_PERIODIC,
                           d3{s.box[0], s.box[1], s.box[2]}, ctx->force_fx.as<u64>(), ctx->energy_fx.as<u64>() + 0,
                           one && G ? ctx->g_eu1.as<double>() : nullptr, E->q.as<double>(), E->dq.as<double>());
        if (one && G)
            launch_group_sums(st, G, ctx->g_wstart.as<int>(), 1, nullptr, ctx->g_eu1.as<double>(),
                              ctx->g_energy.as<u64>(), 0);
    }
    if (int e = build_rows(ctx)) return e;
    const ElecRows R = rows(E);
    // fixed field: real-space part at the polarisable sites (with the rows,
    // build_rows), plus the reciprocal part at every site
    if (P.pme) {
        if (int e = pme_grid(ctx, 0, nullptr)) return e;
        launch_gather(st, n, pos, P, E->grid.as<double>(), E->phi_fix.as<double>());
        launch_record_field(st, n, P, E->phi_fix.as<double>(), 1, 0.0, nullptr, E->efix.as<double>());
    }
    // alpha * E_fixed, and the initial guess (into mu_a)
    launch_scale_copy(st, n, E->efix.as<double>(), E->pol.as<double>(), E->efix_pol.as<double>(), nullptr);
    double* mu_in = E->mu_a.as<double>();
    double* mu_out = E->mu_b.as<double>();
    const size_t mbytes = sizeof(double) * 3 * n;
    switch (E->guess_mode) {
        case MBPOL_GUESS_LEARNED:
            if (int e = guess_predict(ctx, mu_in)) return e;
            break;
        case MBPOL_GUESS_ZERO: launch_fill_zero(st, mu_in, mbytes); break;
        case MBPOL_GUESS_PREVIOUS:
            if (E->have_previous) {
                HIP_TRY(hipMemcpyAsync(mu_in, E->mu_prev.p, mbytes, hipMemcpyDeviceToDevice, st));
                break;
            }
            [[fallthrough]];
        case MBPOL_GUESS_USER:
            if (E->guess_mode == MBPOL_GUESS_USER) {
                HIP_TRY(hipMemcpyAsync(mu_in, E->mu_user.p, mbytes, hipMemcpyDeviceToDevice, st));
                break;
            }
            [[fallthrough]];
        default: HIP_TRY(hipMemcpyAsync(mu_in, E->efix_pol.p, mbytes, hipMemcpyDeviceToDevice, st));
    }
    timer_stop(ctx, T_ELEC_SETUP);
    // the caller's work for the other stream is enqueued here: the set-up above
    // keeps the GPU busy meanwhile, and the solve below waits on the host
    if (after_setup)
        if (int e = after_setup()) return e;
    const double self_term = P.pme ? (4.0 / 3.0) * (P.alpha * P.alpha * P.alpha) / mbpol_params::el_sqrt_pi : 0.0;
    SolveCtl c{};
    if (fixed_steps > 0 && (E->solver != MBPOL_SOLVER_PCG || E->history))
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "a fixed-length solve is a conjugate-gradient one without history");
    if (E->solver == MBPOL_SOLVER_PCG) {
        // conjugate gradients from the guess in mu_in; the solution stays in mu_in, and
        // the reciprocal induced potential of the energy is taken from it (mu_out = mu_in)
        if (int e = solve_pcg(ctx, mu_in, self_term, c, fixed_steps)) return e;
        mu_out = mu_in;
    } else {
        if (E->ngroups)
            return set_error(MBPOL_ERR_INVALID_ARGUMENT, "independent groups are solved by conjugate gradients only");
        if (int e = solve_sor(ctx, E->mu_a.as<double>(), E->mu_b.as<double>(), self_term, mu_in, mu_out, c)) return e;
    }
    E->mu_cur = mu_in;
    if (fixed_steps <= 0) {
        const int it = c.iterations;
        const double cur = c.eps;
        const bool conv = c.converged != 0;
        E->last_iters = it;
        ctx->stats.iterations = it;
        ctx->stats.epsilon = cur;
        ctx->stats.converged = conv ? 1 : 0;
        if (!conv)
            return set_error(MBPOL_ERR_NOT_CONVERGED, "Induced dipoles did not converge:  iterations=%d eps=%g", it,
                             cur);
    }
    // energy and forces (calculateElectrostatic, :838-855 direct sum, :2718-2777 PME)
    timer_start(ctx, T_ELEC_ENERGY);
    launch_elec_energy(st, n, pos, S, P, R, mu_in, E->fsite.as<double>(), E->V.as<double>(), E->eblk.as<double>(),
                       E->ngroups ? E->g_esite.as<double>() : nullptr);
    if (E->ngroups)
        launch_group_sums(st, E->ngroups, ctx->g_wstart.as<int>(), 4, nullptr, E->g_esite.as<double>(),
                          ctx->g_energy.as<u64>(), 4);
    launch_energy_finish(st, energy_blocks(n), E->eblk.as<double>(), energy);
    if (P.pme) {
        // the reciprocal induced potential of the reference is the one of the last
        // iteration's field evaluation, i.e. of the dipoles before the final step
        if (int e = pme_grid(ctx, 1, mu_out)) return e;
        launch_gather(st, n, pos, P, E->grid.as<double>(), E->phi_ind.as<double>());
        launch_recip_site(st, n, S, P, E->phi_fix.as<double>(), E->phi_ind.as<double>(), mu_in, E->fsite.as<double>(),
                          E->V.as<double>(), E->eblk.as<double>());
        launch_energy_finish(st, sor_blocks(n), E->eblk.as<double>(), energy);
    }
    if (P.redistribute) launch_charge_deriv_forces(st, E->nw, E->dq.as<double>(), E->V.as<double>(), E->fsite.as<double>());
    launch_add_forces_fix(st, n, E->fsite.as<double>(), ctx->force_fx.as<u64>());
    timer_stop(ctx, T_ELEC_ENERGY);
    // history of the learned previous-step guess: what was the last evaluation's becomes
    // the one before it
    if (E->history) {
        if (E->have_previous) {
            HIP_TRY(hipMemcpyAsync(E->mu_prev2.p, E->mu_prev.p, mbytes, hipMemcpyDeviceToDevice, st));
            HIP_TRY(hipMemcpyAsync(E->efix_prev2.p, E->efix_prev.p, mbytes, hipMemcpyDeviceToDevice, st));
            E->have_previous2 = true;
        }
        HIP_TRY(hipMemcpyAsync(E->efix_prev.p, E->efix_pol.p, mbytes, hipMemcpyDeviceToDevice, st));
    }
    // converged dipoles kept for the next evaluation's "previous" start (a
    // fixed-length solve sets the flag once its convergence has been read)
    if (fixed_steps <= 0) {
        HIP_TRY(hipMemcpyAsync(E->mu_prev.p, mu_in, mbytes, hipMemcpyDeviceToDevice, st));
        E->have_previous = true;
    } else {
        launch_copy_if_converged(st, 3LL * n, E->ctl.as<SolveCtl>(), mu_in, E->mu_prev.as<double>());
    }
    return MBPOL_OK;
}

}  // namespace mbpol

using mbpol::Elec;

extern "C" {

int mbpol_context_get_dipoles(const mbpol_context* ctx, double* dipoles) {
    using namespace mbpol;
    if (!ctx->elec) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "no electrostatics in this system");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(copy_on(ctx->stream, dipoles, ctx->elec->mu_cur, sizeof(double) * 3 * ctx->elec->n));
    return MBPOL_OK;
}

int mbpol_context_get_charges(const mbpol_context* ctx, double* charges) {
    using namespace mbpol;
    if (!ctx->elec) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "no electrostatics in this system");
    HIP_TRY(copy_on(ctx->stream, charges, ctx->elec->q.p, sizeof(double) * ctx->elec->n));
    return MBPOL_OK;
}

int mbpol_context_set_dipole_guess(mbpol_context* ctx, int32_t mode, const double* dipoles) {
    using namespace mbpol;
    if (ctx) ++ctx->graph_gen;
    if (!ctx->elec) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "no electrostatics in this system");
    if (mode < MBPOL_GUESS_ALPHA_E || mode > MBPOL_GUESS_LEARNED)
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "bad guess mode");
    if (mode == MBPOL_GUESS_LEARNED && !ctx->guess)
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "no learned guess attached (mbpol_guess_attach)");
    if (mode == MBPOL_GUESS_USER) {
        if (!dipoles) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "user guess needs dipoles");
        // sites of zero polarisability keep zero dipoles (their guess entries are ignored)
        Elec* E = ctx->elec;
        std::vector<double> g(dipoles, dipoles + 3 * (size_t)E->n);
        for (int i = 0; i < E->n; ++i)
            if (!E->h_pol[i]) g[3 * i] = g[3 * i + 1] = g[3 * i + 2] = 0.0;
        HIP_TRY(hipSetDevice(ctx->device));
        HIP_TRY(copy_on(ctx->stream, E->mu_user.p, g.data(), sizeof(double) * 3 * (size_t)E->n));
    }
    ctx->elec->guess_mode = mode;
    return MBPOL_OK;
}

int mbpol_context_set_dipole_convergence(mbpol_context* ctx, double target_eps, int32_t max_iter) {
    using namespace mbpol;
    if (ctx) {
        ++ctx->graph_gen;
        ctx->graph_steps = 0;  // the captured solve's length is taken anew from the next solve
    }
    if (!ctx) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null context");
    if (!ctx->elec) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "no electrostatics in this system");
    if (!(target_eps > 0.0) || max_iter < 1)
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "target epsilon must be > 0 and max iterations >= 1");
    ctx->sys.el_eps = target_eps;
    ctx->sys.el_max_iter = max_iter;
    return MBPOL_OK;
}

int mbpol_context_get_tuning(const mbpol_context* ctx, const char* name, int32_t* value) {
    using namespace mbpol;
    if (!ctx || !name || !value) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null argument");
    // launch shapes and capacities of the whole evaluation, and the graph's counters
    const struct {
        const char* name;
        long long v;
    } own[] = {{"fork_early", ctx->fork_early},
               {"pip_waves", ctx->pip_waves},
               {"pip3_waves", ctx->pip3_waves},
               {"disp_waves", ctx->disp_waves},
               {"tune_evaluations", ctx->tune_evaluations},
               {"pip_blocks", ctx->pip_blocks},
               {"pip3_blocks", ctx->pip3_blocks},
               {"pip3_kernel", ctx->pip3_kernel},
               {"pip2_kernel", ctx->pip2_kernel},
               {"disp_blocks", ctx->disp_blocks},
               {"dimer_capacity", ctx->dimer_cap},
               {"lower_capacity", ctx->lower_cap},
               {"trimer_capacity", ctx->trip_cap},
               {"graph_steps", ctx->graph_steps},
               {"graph_captures", ctx->graph_captures},
               {"graph_replays", ctx->graph_replays},
               {"graph_fallbacks", ctx->graph_fallbacks}};
    for (const auto& o : own)
        if (std::strcmp(name, o.name) == 0) {
            *value = (int32_t)std::min<long long>(o.v, INT32_MAX);
            return MBPOL_OK;
        }
    int v = 0;
    if (int e = elec_tuning(ctx, name, &v)) return e;
    *value = v;
    return MBPOL_OK;
}

int mbpol_context_set_solver(mbpol_context* ctx, int32_t solver) {
    using namespace mbpol;
    if (ctx) ++ctx->graph_gen;
    if (!ctx) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null context");
    if (!ctx->elec) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "no electrostatics in this system");
    if (solver != MBPOL_SOLVER_SOR && solver != MBPOL_SOLVER_PCG)
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "unknown dipole solver %d", solver);
    ctx->elec->solver = solver;
    return MBPOL_OK;
}

int mbpol_context_get_electrostatic_potential(mbpol_context* ctx, int32_t npoints, const double* points,
                                              double* potential) {
    using namespace mbpol;
    if (!ctx->elec) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "no electrostatics in this system");
    if (npoints <= 0) return MBPOL_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    DevBuf pts, out;
    HIP_TRY(pts.ensure(sizeof(double) * 3 * npoints));
    HIP_TRY(out.ensure(sizeof(double) * npoints));
    HIP_TRY(copy_on(ctx->stream, pts.p, points, sizeof(double) * 3 * npoints));
    Elec* E = ctx->elec;
    launch_potential_at_points(ctx->stream, npoints, pts.as<double>(), E->n, ctx->pos.as<double>(), E->q.as<double>(),
                               E->mu_cur, E->P, out.as<double>());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(copy_on(ctx->stream, potential, out.p, sizeof(double) * npoints));
    return MBPOL_OK;
}

int mbpol_context_get_electrostatic_potential_noalloc(mbpol_context* ctx, int32_t npoints, const double* points,
                                                      double* potential) {
    using namespace mbpol;
    if (!ctx || (npoints > 0 && (!points || !potential))) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->elec) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "no electrostatics in this system");
    if (npoints < 0 || npoints > kScratchPoints)
        return set_error(MBPOL_ERR_INVALID_ARGUMENT, "potential without allocation: at most %d points per call (%d asked)",
                         kScratchPoints, npoints);
    if (npoints == 0) return MBPOL_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    double* sc = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        sc = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch_ptr[ctx->device] : nullptr;
    }
    if (!sc) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "scratch not set up on device %d", ctx->device);
    double* pts = sc;
    double* out = sc + 3 * (size_t)kScratchPoints;
    Elec* E = ctx->elec;
    HIP_TRY(hipMemcpyAsync(pts, points, sizeof(double) * 3 * npoints, hipMemcpyHostToDevice, ctx->stream));
    launch_potential_at_points(ctx->stream, npoints, pts, E->n, ctx->pos.as<double>(), E->q.as<double>(), E->mu_cur,
                               E->P, out);
    HIP_TRY(hipMemcpyAsync(potential, out, sizeof(double) * npoints, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MBPOL_OK;
}

int mbpol_context_get_system_electrostatics_moments(mbpol_context* ctx, const double* masses, double* moments) {
    using namespace mbpol;
    if (!ctx || !masses || !moments) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->elec) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "no electrostatics in this system");
    HIP_TRY(hipSetDevice(ctx->device));
    Elec* E = ctx->elec;
    DevBuf m, out;
    HIP_TRY(m.ensure(sizeof(double) * E->n));
    HIP_TRY(out.ensure(sizeof(double) * 13));
    HIP_TRY(copy_on(ctx->stream, m.p, masses, sizeof(double) * E->n));
    launch_system_moments(ctx->stream, E->n, ctx->pos.as<double>(), E->q.as<double>(), E->mu_cur, m.as<double>(),
                          out.as<double>());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(copy_on(ctx->stream, moments, out.p, sizeof(double) * 13));
    return MBPOL_OK;
}

int mbpol_context_update_electrostatics(mbpol_context* ctx, int32_t n, const double* charge, const int32_t* molecule,
                                        const int32_t* atom_type, const double* damping, const double* polarity) {
    using namespace mbpol;
    ++ctx->graph_gen;
    if (!ctx->elec) return set_error(MBPOL_ERR_INVALID_ARGUMENT, "no electrostatics in this system");
    if (n != ctx->elec->n)
        return set_error(MBPOL_ERR_COUNT_CHANGED, "updateParametersInContext: The number of multipoles has changed");
    mbpol_system& s = ctx->sys;
    s.el_q.assign(charge, charge + n);
    s.el_mol.assign(molecule, molecule + n);
    s.el_type.assign(atom_type, atom_type + n);
    s.el_damp.assign(damping, damping + n);
    s.el_pol.assign(polarity, polarity + n);
    HIP_TRY(hipSetDevice(ctx->device));
    Elec* E = ctx->elec;
    if (int e = up(ctx->stream, E->q0, s.el_q)) return e;
    if (int e = up(ctx->stream, E->damp6, sixth_roots(s.el_damp))) return e;
    if (int e = up(ctx->stream, E->pol, s.el_pol)) return e;
    if (int e = up(ctx->stream, E->mol, s.el_mol)) return e;
    if (int e = up(ctx->stream, E->type, s.el_type)) return e;
    if (int e = set_polarisable(ctx, E, s.el_pol, s.el_mol)) return e;
    HIP_TRY(E->partial.ensure(sizeof(double) * (iterate_blocks(E->npol) + 1)));
    // the polarisable set may have changed: clear the dipoles the iteration no longer visits
    launch_fill_zero(ctx->stream, E->mu_a.p, sizeof(double) * 3 * std::max(n, 1));
    launch_fill_zero(ctx->stream, E->mu_b.p, sizeof(double) * 3 * std::max(n, 1));
    launch_fill_zero(ctx->stream, E->mu_prev.p, sizeof(double) * 3 * std::max(n, 1));
    E->mu_cur = E->mu_a.as<double>();
    E->have_previous = false;
    E->have_previous2 = false;
    if (int e = set_groups(ctx)) return e;
    if (!E->P.pme) set_direct_caps(E);
    return MBPOL_OK;
}

}  // extern "C"
