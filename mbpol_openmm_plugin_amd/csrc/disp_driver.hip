// Host driver of the damped-C6 dispersion (python/mbpol.xml:37-108).
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/mbpol_hip.h"
#include "disp_kernels.h"
#include "elec.hpp"
#include "elec_kernels.h"

namespace mbpol {

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t _e = (expr);                                                                         \
        if (_e != hipSuccess) return set_error(MBPOL_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

struct Disp {
    DispGeom g{};
    DevBuf act, type, mol, c6, d6, cell_of, cell_count, cell_start, cell_fill, members, unsorted, eblk, npairs;
    DevBuf cnt;  // pairs within the cutoff per active atom
    DevBuf gs, ge;  // independent groups: candidate range (in active-atom order) of every active atom
    DevBuf gact;    // independent groups: first active atom of every group (ngroups + 1)
    DevBuf cpos;  // positions, molecule and type of the active atoms in cell order
    int ncell = 1;
};

template <typename T>
static int upv(hipStream_t stream, DevBuf& b, const std::vector<T>& v) {
    HIP_TRY(b.ensure(std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty()) HIP_TRY(copy_on(stream, b.p, v.data(), v.size() * sizeof(T)));
    return MBPOL_OK;
}

int disp_set_box(mbpol_context* ctx) {
    Disp* D = ctx->disp;
    const mbpol_system& s = ctx->sys;
    D->g.L = d3{s.box[0], s.box[1], s.box[2]};
    D->g.periodic = s.ds_method == MBPOL_CUTOFF_PERIODIC;
    D->g.use_cells = 0;
    D->g.nc = make_int3(1, 1, 1);
    if (D->g.periodic) {
        int nc[3];
        for (int d = 0; d < 3; ++d) nc[d] = (int)std::floor(s.box[d] / s.ds_cutoff);
        if (nc[0] >= 3 && nc[1] >= 3 && nc[2] >= 3) {
            D->g.use_cells = 1;
            D->g.nc = make_int3(nc[0], nc[1], nc[2]);
        }
    }
    D->ncell = D->g.nc.x * D->g.nc.y * D->g.nc.z;
    HIP_TRY(D->cell_count.ensure(sizeof(int) * (D->ncell + 1)));
    HIP_TRY(D->cell_start.ensure(sizeof(int) * (D->ncell + 1)));
    HIP_TRY(D->cell_fill.ensure(sizeof(int) * (D->ncell + 1)));
    return MBPOL_OK;
}

int disp_create(mbpol_context* ctx) {
    const mbpol_system& s = ctx->sys;
    Disp* D = new Disp();
    ctx->disp = D;
    // particles whose type has a non-zero C6 row take part (M sites do not)
    std::vector<char> active(s.ds_ntypes, 0);
    for (int t = 0; t < s.ds_ntypes; ++t)
        for (int u = 0; u < s.ds_ntypes; ++u)
            if (s.ds_c6[t + s.ds_ntypes * u] != 0.0) active[t] = 1;
    std::vector<int> act;
    for (int i = 0; i < s.nd; ++i)
        if (active[s.ds_type[i]]) act.push_back(i);
    D->g.na = (int)act.size();
    D->g.ntypes = s.ds_ntypes;
    D->g.method = s.ds_method;
    D->g.cutoff = s.ds_cutoff;
    if (int e = upv(ctx->stream, D->act, act)) return e;
    if (int e = upv(ctx->stream, D->type, s.ds_type)) return e;
    if (int e = upv(ctx->stream, D->mol, s.ds_mol)) return e;
    if (int e = upv(ctx->stream, D->c6, s.ds_c6)) return e;
    if (int e = upv(ctx->stream, D->d6, s.ds_d6)) return e;
    HIP_TRY(D->cell_of.ensure(sizeof(int) * std::max(1, D->g.na)));
    HIP_TRY(D->members.ensure(sizeof(int) * std::max(1, D->g.na)));
    HIP_TRY(D->cpos.ensure(sizeof(double4) * std::max(1, D->g.na)));
    HIP_TRY(D->unsorted.ensure(sizeof(int) * std::max(1, D->g.na)));
    HIP_TRY(D->eblk.ensure(sizeof(double) * std::max(1, D->g.na)));
    HIP_TRY(D->npairs.ensure(sizeof(unsigned long long)));
    HIP_TRY(D->cnt.ensure(sizeof(unsigned long long) * std::max(1, D->g.na)));
    if (!s.grp_wstart.empty()) {
        // active atoms are in particle order, so every group's atoms are consecutive
        const std::vector<int>& ws = s.grp_wstart;
        const int G = (int)ws.size() - 1;
        std::vector<int> gact(G + 1, 0), gs(D->g.na), ge(D->g.na);
        for (int g = 0, a = 0; g <= G; ++g) {
            while (a < D->g.na && g < G && act[a] / 4 < ws[g]) ++a;
            gact[g] = g < G ? a : D->g.na;
        }
        for (int g = 0; g < G; ++g)
            for (int a = gact[g]; a < gact[g + 1]; ++a) {
                gs[a] = gact[g];
                ge[a] = gact[g + 1];
            }
        if (int e = upv(ctx->stream, D->gs, gs)) return e;
        if (int e = upv(ctx->stream, D->ge, ge)) return e;
        if (int e = upv(ctx->stream, D->gact, gact)) return e;
        D->g.gs = D->gs.as<int>();
        D->g.ge = D->ge.as<int>();
    }
    return disp_set_box(ctx);
}

int disp_group_sums(mbpol_context* ctx, hipStream_t st, u64* out) {
    const Disp* D = ctx->disp;
    if (!D->g.gs) return MBPOL_OK;
    // eblk[a] holds half of atom a's pair energies; every pair is seen from both atoms
    launch_group_sums(st, ctx->ngroups, D->gact.as<int>(), 1, nullptr, D->eblk.as<double>(), out, 3);
    return MBPOL_OK;
}

void disp_destroy(mbpol_context* ctx) {
    delete ctx->disp;
    ctx->disp = nullptr;
}

void launch_energy_finish(hipStream_t s, int nblocks, const double* eblk, u64* energy);

int disp_evaluate(mbpol_context* ctx, hipStream_t st, u64* energy) {
    Disp* D = ctx->disp;
    const double* pos = ctx->pos.as<double>();
    if (D->g.na == 0) return MBPOL_OK;
    if (D->g.use_cells) {
        launch_fill_zero(st, D->cell_count.p, sizeof(int) * (D->ncell + 1));
        launch_fill_zero(st, D->cell_fill.p, sizeof(int) * (D->ncell + 1));
        launch_disp_cells(st, D->g.na, D->act.as<int>(), pos, D->g.L, D->g.nc, D->cell_of.as<int>(), D->cell_count.as<int>());
        if (int e = scan_i32(ctx, D->cell_count.as<int>(), D->cell_start.as<int>(), D->ncell + 1, st)) return e;
        launch_disp_cell_fill(st, D->g.na, D->cell_of.as<int>(), D->cell_start.as<int>(), D->cell_fill.as<int>(),
                              D->unsorted.as<int>());
        // members in index order: the pair sums, and so the forces, do not depend on scheduling
        launch_rank_sort(st, D->g.na, D->cell_of.as<int>(), D->cell_start.as<int>(), D->unsorted.as<int>(),
                         D->members.as<int>());
        launch_disp_pack(st, D->g.na, D->members.as<int>(), D->act.as<int>(), pos, D->type.as<int>(),
                         D->mol.as<int>(), D->cpos.as<double4>());
    }
    launch_disp(st, ctx->precision == MBPOL_PRECISION_DOUBLE, D->g, D->act.as<int>(), pos, D->type.as<int>(),
                D->mol.as<int>(), D->c6.as<double>(), D->d6.as<double>(), D->cell_of.as<int>(), D->cell_start.as<int>(),
                D->members.as<int>(), D->cpos.as<double4>(), ctx->force_fx.as<u64>(), D->eblk.as<double>(),
                D->cnt.as<unsigned long long>(), ctx->disp_blocks);
    launch_count_total(st, D->g.na, D->cnt.as<unsigned long long>(), D->npairs.as<unsigned long long>());
    launch_energy_finish(st, D->g.na, D->eblk.as<double>(), energy);
    return MBPOL_OK;
}

int disp_pair_count(const mbpol_context* ctx, long long* pairs) {
    const Disp* D = ctx->disp;
    *pairs = 0;
    if (!D || D->g.na == 0) return MBPOL_OK;
    unsigned long long v = 0;
    HIP_TRY(copy_on(ctx->stream, &v, D->npairs.p, sizeof(v)));
    *pairs = (long long)(v / 2);  // each pair is counted from both atoms
    return MBPOL_OK;
}

}  // namespace mbpol
