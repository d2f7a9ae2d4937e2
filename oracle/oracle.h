/* C ABI of the FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY (tests/, smoke(),
 * bench.py's cpu_baseline).  Units at this boundary follow the reference's
 * OpenMM API: positions nm, energies kJ/mol, forces kJ/mol/nm (accumulated). */
#ifndef MBPOL_ORACLE_H
#define MBPOL_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int method;              /* 0 = NoCutoff, 1 = PME */
    double cutoff;           /* nm, PME real space */
    double alpha;            /* nm^-1 */
    int grid[3];
    double thole[5];         /* TCC, TCD, TDD, TDDOH, TDDHH */
    int max_iter;
    double target_eps;
    int charge_redistribution;
    double box[3];           /* nm */
} oracle_elec_params;

typedef struct {
    int iterations;
    double epsilon;
    int converged;
    /* wall-clock seconds of the phases (for the CPU throughput estimate of a
     * sampled run): charges + fixed field + pair scales, one dipole iteration
     * on average, energy/forces */
    double setup_s, iteration_s, energy_s;
} oracle_elec_stats;

/* polynomial kernels */
double oracle_pip2b(const double* x31, double* g31);
double oracle_pip3b(const double* x36, double* g36);

/* components; idx = per-molecule {O,H1,H2} particle indices */
double oracle_onebody(int nw, const double* pos, const int* idx, int periodic, const double* box, double* forces);
int64_t oracle_dimer_list(int n, const double* opos, const double* box, int periodic, double rc, int* out, int64_t max_out);
int64_t oracle_trimer_list(int n, const double* opos, const double* box, int periodic, double rc, double rmin, int* out,
                           int64_t max_out);
double oracle_twobody(int npairs, const int* pairs, const double* pos, const int* idx, int periodic, const double* box,
                      double* forces);
double oracle_threebody(int64_t ntrip, const int* trip, const double* pos, const int* idx, int periodic, const double* box,
                        double* forces);
/* electrostatics; returns 0 on success, 1 when the dipoles did not converge */
int oracle_electrostatics(const oracle_elec_params* p, int n, const double* pos, const double* charge, const int* mol,
                          const int* type, const double* damp, const double* pol, double* forces, double* energy,
                          oracle_elec_stats* stats, double* dipoles_out, double* fixed_field_out, char* err, int errlen);
/* getElectrostaticPotential: potential (kJ/mol/e) at npts points (nm) from the site charges and converged
 * induced dipoles; returns 0, or 1 when the dipoles did not converge */
int oracle_electrostatic_potential(const oracle_elec_params* p, int n, const double* pos, const double* charge,
                                   const int* mol, const int* type, const double* damp, const double* pol, int npts,
                                   const double* pts, double* out);
/* getSystemElectrostaticsMoments: out[13] = net charge, dipole (3, Debye), traceless quadrupole (3x3, Buckingham)
 * about the centre of mass of `masses`; returns 0, or 1 when the dipoles did not converge */
int oracle_system_moments(const oracle_elec_params* p, int n, const double* pos, const double* charge, const int* mol,
                          const int* type, const double* damp, const double* pol, const double* masses, double* out);
double oracle_thole_scale(const oracle_elec_params* p, double damp_i, double damp_j, int mol_i, int mol_j, int type_i,
                          int type_j, double r, int just_scale, int order, int itype);
void oracle_water_charges(const double* pos4, double* q4, double* dq);
int oracle_solve_given_field(const oracle_elec_params* p, int n, const double* pos, const int* mol, const int* type,
                             const double* damp, const double* pol, const double* field, double* dipoles,
                             oracle_elec_stats* stats);
double oracle_gammq(double a, double x);
/* damped C6 dispersion over particles with types; pairs in the same molecule
 * excluded; types with an all-zero C6 row skipped; method 0 = no cutoff,
 * 1 = cutoff periodic, 2 = cutoff non-periodic */
double oracle_dispersion(int n, const double* pos, const int* type, const int* mol, int ntypes, const double* c6,
                         const double* d6, int method, double cutoff, const double* box, double* forces);
/* FFT self-check helper: in-place 3D complex transform, sign -1 forward */
void oracle_fft3d(double* grid_interleaved, int nx, int ny, int nz, int sign);

#ifdef __cplusplus
}
#endif
#endif
