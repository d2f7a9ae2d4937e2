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
    int redistribute;
};

struct SiteParams {
    const double* q;
    const double* damp;
    const double* pol;
    const int* mol;
    const int* type;
};

void launch_site_charges_copy(hipStream_t s, int n, const double* q0, double* q);
void launch_site_cells(hipStream_t s, int n, const double* pos, ElecParams P, int3 nc, int* cell_of, int* cell_count);
void launch_site_cell_fill(hipStream_t s, int n, const int* cell_of, const int* cell_start, int* cell_fill, int* members);
void launch_site_pairs(hipStream_t s, int n, const double* pos, ElecParams P, int3 nc, int use_cells, const int* cell_of,
                       const int* cell_start, const int* members, int* count, const int* start, int* out);
void launch_pair_prefactors(hipStream_t s, int n, const double* pos, SiteParams S, ElecParams P, const int* nb_start,
                            const int* nb, double2* pf);
void launch_fixed_field(hipStream_t s, int n, const double* pos, SiteParams S, ElecParams P, const int* nb_start,
                        const int* nb, double* field);
void launch_induced_field(hipStream_t s, int n, const double* pos, ElecParams P, const int* nb_start, const int* nb,
                          const double2* pf, const double* mu, double* field);
void launch_spread(hipStream_t s, int n, int mode, const double* pos, const double* q, const double* mu, ElecParams P,
                   double* grid);
void launch_convolve(hipStream_t s, size_t nhalf, const double* influence, double2* spec);
void launch_influence(hipStream_t s, ElecParams P, const double* bx, const double* by, const double* bz, double* out);
void launch_gather(hipStream_t s, int n, const double* pos, ElecParams P, const double* grid, double* phi);
void launch_record_field(hipStream_t s, int n, ElecParams P, const double* phi, int accumulate, double self_term,
                         const double* mu, double* field);
int sor_blocks(int n);
void launch_sor_update(hipStream_t s, int n, const double* pol, const double* efix_pol, const double* field, double omega,
                       double* mu, double* partial);
void launch_eps_finish(hipStream_t s, int nblocks, int n, const double* partial, double* eps_out);
void launch_scale_copy(hipStream_t s, int n, const double* a, const double* pol, double* out, double* out2);
int energy_blocks(int n);
void launch_elec_energy(hipStream_t s, int n, const double* pos, SiteParams S, ElecParams P, const int* nb_start,
                        const int* nb, const double* mu, double* force, double* V, double* eblk);
void launch_recip_site(hipStream_t s, int n, SiteParams S, ElecParams P, const double* phi_fix, const double* phi_ind,
                       const double* mu, double* force, double* V, double* eblk);
void launch_charge_deriv_forces(hipStream_t s, int nw, const double* dq, const double* V, double* force);
void launch_energy_finish(hipStream_t s, int nblocks, const double* eblk, u64* energy);
void launch_add_forces_fix(hipStream_t s, int n, const double* f, u64* fx);
void launch_potential_at_points(hipStream_t s, int npts, const double* pts, int n, const double* pos, const double* q,
                                const double* mu, ElecParams P, double* out);

}  // namespace mbpol
