// Types and launchers of the dispersion kernels (kernels_disp.hip).
#pragma once
#include "kernels.h"

namespace mbpol {

struct DispGeom {
    int na;         // active particles (types with a non-zero C6 row)
    int ntypes;
    int method;     // 0 no cutoff, 1 cutoff periodic, 2 cutoff non-periodic
    int periodic;
    double cutoff;  // nm
    d3 L;
    int use_cells;
    int3 nc;
    // independent groups (non-periodic): candidates of active atom a are [gs[a], ge[a]); null = all
    const int *gs, *ge;
};

void launch_disp_cells(hipStream_t s, int na, const int* act, const double* pos, d3 L, int3 nc, int* cell_of,
                       int* cell_count);
void launch_disp_cell_fill(hipStream_t s, int na, const int* cell_of, const int* cell_start, int* cell_fill, int* members);
// cpos[t] = (position, molecule * 16 + type) of atom act[members[t]] (types < 16)
void launch_disp_pack(hipStream_t s, int na, const int* members, const int* act, const double* pos, const int* type,
                      const int* mol, double4* cpos);
void launch_disp(hipStream_t s, int precision_double, DispGeom g, const int* act, const double* pos, const int* type,
                 const int* mol, const double* c6, const double* d6, const int* cell_of, const int* cell_start,
                 const int* members, const double4* cpos, u64* force, double* eblk, unsigned long long* cnt,
                 int max_blocks = 0);  // max_blocks > 0: capped grid, each wave strides over the atoms
// cnt: per active atom; launch_count_total sums it into *out
void launch_count_total(hipStream_t s, int na, const unsigned long long* cnt, unsigned long long* out);

}  // namespace mbpol
