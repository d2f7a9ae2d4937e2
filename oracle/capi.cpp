// C ABI of the oracle (see oracle.h) -- test infrastructure only.
#include <complex>
#include <cstring>
#include <stdexcept>
#include "oracle.h"
#include "oracle_common.h"

namespace oracle {
double onebody_all(int nw, const double* pos, const int* idx, bool periodic, const double* box, double* forces);
int64_t dimer_list(int n, const double* opos, const double* box, bool periodic, double rc, int* out, int64_t max_out);
int64_t trimer_list(int n, const double* opos, const double* box, bool periodic, double rc, double rmin, int* out,
                    int64_t max_out);
double twobody_all(int npairs, const int* pairs, const double* pos, const int* idx, bool periodic, const double* box,
                   double* forces);
double threebody_all(int64_t ntrip, const int* trip, const double* pos, const int* idx, bool periodic, const double* box,
                     double* forces);
double electrostatics(const oracle_elec_params& P, int n, const double* pos, const double* charge, const int* mol,
                      const int* type, const double* damp, const double* pol, double* forces, oracle_elec_stats* stats,
                      double* dipoles_out, double* fixed_field_out);
int electrostatic_potential(const oracle_elec_params& P, int n, const double* pos, const double* charge,
                            const int* mol, const int* type, const double* damp, const double* pol, int npts,
                            const double* pts, double* out);
int system_moments(const oracle_elec_params& P, int n, const double* pos, const double* charge, const int* mol,
                   const int* type, const double* damp, const double* pol, const double* masses, double* out);
double thole_scale(const oracle_elec_params& P, double damp_i, double damp_j, int mol_i, int mol_j, int type_i, int type_j,
                   double r, int just_scale, int order, int itype);
void water_charges(const double* pos4, double* q4, double* dq);
int solve_given_field(const oracle_elec_params& P, int n, const double* pos, const int* mol, const int* type,
                      const double* damp, const double* pol, const double* field, double* dipoles, oracle_elec_stats* stats);
double gammq(double a, double x);
double dispersion(int n, const double* pos, const int* type, const int* mol, int ntypes, const double* c6,
                  const double* d6, int method, double cutoff, const double* box, double* forces);
void fft3d(std::complex<double>* grid, int nx, int ny, int nz, int sign);
}  // namespace oracle

extern "C" {

double oracle_pip2b(const double* x, double* g) { return oracle::pip2b_eval(x, g); }
double oracle_pip3b(const double* x, double* g) { return oracle::pip3b_eval(x, g); }

double oracle_onebody(int nw, const double* pos, const int* idx, int periodic, const double* box, double* forces) {
    return oracle::onebody_all(nw, pos, idx, periodic != 0, box, forces);
}
int64_t oracle_dimer_list(int n, const double* opos, const double* box, int periodic, double rc, int* out, int64_t max_out) {
    return oracle::dimer_list(n, opos, box, periodic != 0, rc, out, max_out);
}
int64_t oracle_trimer_list(int n, const double* opos, const double* box, int periodic, double rc, double rmin, int* out,
                           int64_t max_out) {
    return oracle::trimer_list(n, opos, box, periodic != 0, rc, rmin, out, max_out);
}
double oracle_twobody(int npairs, const int* pairs, const double* pos, const int* idx, int periodic, const double* box,
                      double* forces) {
    return oracle::twobody_all(npairs, pairs, pos, idx, periodic != 0, box, forces);
}
double oracle_threebody(int64_t ntrip, const int* trip, const double* pos, const int* idx, int periodic, const double* box,
                        double* forces) {
    return oracle::threebody_all(ntrip, trip, pos, idx, periodic != 0, box, forces);
}
int oracle_electrostatics(const oracle_elec_params* p, int n, const double* pos, const double* charge, const int* mol,
                          const int* type, const double* damp, const double* pol, double* forces, double* energy,
                          oracle_elec_stats* stats, double* dipoles_out, double* fixed_field_out, char* err, int errlen) {
    try {
        *energy = oracle::electrostatics(*p, n, pos, charge, mol, type, damp, pol, forces, stats, dipoles_out, fixed_field_out);
        return 0;
    } catch (const std::exception& e) {
        if (err && errlen > 0) {
            std::strncpy(err, e.what(), errlen - 1);
            err[errlen - 1] = 0;
        }
        return 1;
    }
}
int oracle_electrostatic_potential(const oracle_elec_params* p, int n, const double* pos, const double* charge,
                                   const int* mol, const int* type, const double* damp, const double* pol, int npts,
                                   const double* pts, double* out) {
    return oracle::electrostatic_potential(*p, n, pos, charge, mol, type, damp, pol, npts, pts, out);
}
int oracle_system_moments(const oracle_elec_params* p, int n, const double* pos, const double* charge, const int* mol,
                          const int* type, const double* damp, const double* pol, const double* masses, double* out) {
    return oracle::system_moments(*p, n, pos, charge, mol, type, damp, pol, masses, out);
}
double oracle_thole_scale(const oracle_elec_params* p, double damp_i, double damp_j, int mol_i, int mol_j, int type_i,
                          int type_j, double r, int just_scale, int order, int itype) {
    return oracle::thole_scale(*p, damp_i, damp_j, mol_i, mol_j, type_i, type_j, r, just_scale, order, itype);
}
void oracle_water_charges(const double* pos4, double* q4, double* dq) { oracle::water_charges(pos4, q4, dq); }
int oracle_solve_given_field(const oracle_elec_params* p, int n, const double* pos, const int* mol, const int* type,
                             const double* damp, const double* pol, const double* field, double* dipoles,
                             oracle_elec_stats* stats) {
    return oracle::solve_given_field(*p, n, pos, mol, type, damp, pol, field, dipoles, stats);
}
double oracle_gammq(double a, double x) { return oracle::gammq(a, x); }
double oracle_dispersion(int n, const double* pos, const int* type, const int* mol, int ntypes, const double* c6,
                         const double* d6, int method, double cutoff, const double* box, double* forces) {
    return oracle::dispersion(n, pos, type, mol, ntypes, c6, d6, method, cutoff, box, forces);
}
void oracle_fft3d(double* g, int nx, int ny, int nz, int sign) {
    oracle::fft3d(reinterpret_cast<std::complex<double>*>(g), nx, ny, nz, sign);
}
}
