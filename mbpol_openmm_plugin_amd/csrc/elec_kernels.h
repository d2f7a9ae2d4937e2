// Types and launchers of the electrostatics kernels (kernels_elec.hip).
#pragma once
#include "kernels.h"

namespace mbpol {

struct ElecParams {
    int n;            // sites
    int pme;          // 0 = direct sum over all pairs, 1 = PME
    double cutoff;    // nm (huge without PME)
    double alpha;     // Ewald alpha, nm^-1
    d3 box, inv_box;
    int3 grid;
    double thole[5];  // TCC, TCD, TDD, TDDOH, TDDHH
    double tcc_q;     // TCC^(1/4)
    int redistribute;
    // independent groups (direct sum only): site i interacts with sites [gs[i], ge[i]); null = all
    const int *gs, *ge;
};

struct SiteParams {
    const double* q;
    const double* damp6;  // damping^(1/6)
    const double* pol;
    const int* mol;
    const int* type;
};

// fixed-capacity neighbour rows: row i of P (Q) holds the polarisable
// (non-polarisable) neighbours of site i; pf holds the induced-field
// prefactors of the P entries of polarisable sites
struct ElecRows {
    int capP, capQ;
    const int *nbP, *cntP, *nbQ, *cntQ;
    const double2* pf;
};

// state of one dipole solve, written on the device by its convergence test;
// every later kernel of the solve returns at once when stop is set
struct SolveCtl {
    int stop;
    int converged;
    int iterations;  // field evaluations so far
    int pad;
    double eps;      // the reference's measure at the last test
};

// PME grid bricks: s[d] points per brick along axis d (every brick, the last
// one included, at least 4 points wide), nb[d] bricks
struct BrickGeom {
    int s[3];
    int nb[3];
};

void launch_site_charges_copy(hipStream_t s, int n, const double* q0, double* q);
// dst = src (n doubles) when ctl records a converged stop, else nothing
void launch_copy_if_converged(hipStream_t s, long long n, const SolveCtl* ctl, const double* src, double* dst);
void launch_site_cells(hipStream_t s, int n, const double* pos, ElecParams P, int3 nc, int* cell_of, int* cell_count);
void launch_bucket_fill(hipStream_t s, int n, const int* bucket_of, const int* bucket_start, int* bucket_fill,
                        int* members);
// members_out: every bucket's items in ascending order (members_in: same buckets, any order)
void launch_rank_sort(hipStream_t s, int n, const int* bucket_of, const int* bucket_start, const int* members_in,
                      int* members_out);
// cpos[t] = (position, polarisable ? 1 : 0) of site members[t]
void launch_site_pack(hipStream_t s, int n, const int* members, const double* pos, const double* pol, double4* cpos);
// lanes per site: 8, 16 or 32 (the rows are the same for every choice); cpos used when use_cells
void launch_site_lists(hipStream_t s, int lanes, int n, const double* pos, const double* pol, ElecParams P, int3 nc,
                       int reach, int use_cells, const int* cell_of, const int* cell_start, const int* members,
                       const double4* cpos, int capP, int capQ, int* nbP, int* cntP, int* nbQ, int* cntQ,
                       int* maxcnt);
// prefactors of the P rows (into pf) and the real-space fixed field added to field, in one pass
void launch_prefactors_field(hipStream_t s, int npol, const int* polidx, const double* pos, SiteParams S, ElecParams P,
                             const ElecRows& R, double2* pf, double* field);
void launch_site_bins(hipStream_t s, int n, const double* pos, const double* weight, ElecParams P, BrickGeom B,
                      int* bin_of, int* bin_count);
void launch_site_spline_table(hipStream_t s, int n, const double* pos, ElecParams P, int* g0tab, double* sptab);
void launch_spread_bricks(hipStream_t s, int mode, const double* q, const double* mu, ElecParams P, BrickGeom B,
                          const int* bin_start, const int* members, const int* g0tab, const double* sptab,
                          double* grid);
void launch_csr_count(hipStream_t s, int npol, const int* polidx, const int* g0tab, ElecParams P, int* rowcount);
void launch_csr_fill(hipStream_t s, int npol, const int* polidx, const int* g0tab, const double* sptab, ElecParams P,
                     const int* rowstart, int* fill, int* esite, double* ecoef);
void launch_csr_sort(hipStream_t s, size_t ngrid, const int* rowstart, int* esite, double* ecoef);
void launch_spread_csr(hipStream_t s, size_t ngrid, const int* rowstart, const int* esite, const double* ecoef,
                       const double* mu, double* grid, const KEvents& ev, const int* stop = nullptr);
void launch_convolve(hipStream_t s, size_t nhalf, const double* influence, double2* spec);
// influence function on the half grid, rows of khp entries (entries kz >= K3/2 + 1 are zero)
void launch_influence(hipStream_t s, ElecParams P, int khp, const double* bx, const double* by, const double* bz,
                      double* out);
void launch_gather(hipStream_t s, int n, const double* pos, ElecParams P, const double* grid, double* phi);
void launch_record_field(hipStream_t s, int n, ElecParams P, const double* phi, int accumulate, double self_term,
                         const double* mu, double* field);
int iterate_blocks(int npol);
void launch_iterate(hipStream_t s, int npol, const int* polidx, const double* pos, ElecParams P, const ElecRows& R,
                    const double* mu_in, double* mu_out, const double* efix_pol, const double* pol, const double* grid,
                    double self_term, double omega, double* partial, const SolveCtl* ctl, const KEvents& ev);
// independent groups in one solve: the convergence measure is that of the worst group, from the
// |alpha r|^2 of every preconditioner block (blocks of group g: [blk_start[g], blk_start[g+1]))
struct GroupConv {
    int ngroups;  // 0: one system
    const int* blk_start;
    const int* nsites;  // sites of each group (all of them, as the reference's measure counts)
    const double* blk_ee;
};

// preconditioned conjugate gradients on A mu = E_fix, A = diag(1/alpha) - T (see kernels_elec.hip)
// preconditioner blocks: the polarisable sites of one molecule (at most 4 per block)
struct PrecondBlocks {
    int n;             // blocks
    const int* start;  // n + 1 offsets into site
    const int* site;   // site indices, block by block
};
int cg_blocks(int nitems);
int precond_wgs(int nblocks);  // workgroups (and partial pairs) of launch_cg_precond / launch_cg_update
int block_ld();  // leading dimension of a block's inverse (block_ld()^2 doubles per block)
void launch_block_factor(hipStream_t s, const PrecondBlocks& B, const double* pos, SiteParams S, ElecParams P,
                         double* minv);
// mode 0: out = E_fix - A v; mode 1: out = A v with partials (v.out), 2 per iterate_blocks(npol)
void launch_cg_matvec(hipStream_t s, int npol, const int* polidx, const double* pos, ElecParams P, const ElecRows& R,
                      const double* v, const double* grid, double self_term, const double* pol, const double* efix,
                      int mode, double* out, double* partial, const SolveCtl* ctl, const KEvents& ev);
// z = M^-1 r with partials (r.z, |alpha r|^2), 2 per precond_wgs(B.n)
// blk_ee (optional): |alpha r|^2 of every block
void launch_cg_precond(hipStream_t s, const PrecondBlocks& B, const double* minv, const double* pol, const double* r,
                       double* z, double* partial, double* blk_ee, const SolveCtl* ctl);
// step k (uses rz[k] and the matvec partials); partials as launch_cg_precond
void launch_cg_update(hipStream_t s, const PrecondBlocks& B, const double* minv, const double* pol, int npol,
                      const double* partial_mv, const double* rz, int k, double* x, double* r, const double* p,
                      const double* q, double* z, double* partial, double* blk_ee, const SolveCtl* ctl);
// direction k (writes rz[k], reads rz[k-1]) and the convergence measure from nb block partial pairs
void launch_cg_direction(hipStream_t s, int npol, const int* polidx, int nb, const double* partial, double* rz, int k,
                         const double* z, double* p, int n_all, double tol, int max_iter, const GroupConv& gc,
                         SolveCtl* ctl);
// SOR step t: the measure and the stopping rules into ctl
void launch_sor_finish(hipStream_t s, int nblocks, int n_all, const double* partial, int t, double tol, int max_iter,
                       SolveCtl* ctl);
void launch_scale_copy(hipStream_t s, int n, const double* a, const double* pol, double* out, double* out2);
int energy_blocks(int n);
int sor_blocks(int n);
// esite (optional): energy of every site's pairs (half of each pair), for per-group sums
void launch_elec_energy(hipStream_t s, int n, const double* pos, SiteParams S, ElecParams P, const ElecRows& R,
                        const double* mu, double* force, double* V, double* eblk, double* esite = nullptr);
void launch_recip_site(hipStream_t s, int n, SiteParams S, ElecParams P, const double* phi_fix, const double* phi_ind,
                       const double* mu, double* force, double* V, double* eblk);
void launch_charge_deriv_forces(hipStream_t s, int nw, const double* dq, const double* V, double* force);
void launch_energy_finish(hipStream_t s, int nblocks, const double* eblk, u64* energy);
void launch_add_forces_fix(hipStream_t s, int n, const double* f, u64* fx);
void launch_potential_at_points(hipStream_t s, int npts, const double* pts, int n, const double* pos, const double* q,
                                const double* mu, ElecParams P, double* out);
// system multipole moments about the centre of mass of `mass` (out[13], one block)
void launch_system_moments(hipStream_t s, int n, const double* pos, const double* q, const double* mu,
                           const double* mass, double* out);

}  // namespace mbpol
