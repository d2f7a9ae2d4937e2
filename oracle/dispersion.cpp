// Damped C6 dispersion -- oracle (test infrastructure only).
// Restates the CustomNonbondedForce built by python/mbpol.xml:37-108:
//   E = -C6 * tt6 / r^6,  tt6 = 1 - exp(-d6 r) * sum_{k=0..6} (d6 r)^k / k!
// with C6, d6 from per-type tables (:52-83), same-residue and M pairs
// excluded (:100-104), cutoff truncation without switching.
#include "oracle_common.h"

namespace oracle {

static double periodic_diff(double a, double b, double period) {
    double diff = a - b;
    return diff - std::floor(diff / period + 0.5) * period;
}

double dispersion(int n, const double* pos, const int* type, const int* mol, int ntypes, const double* c6,
                  const double* d6, int method, double cutoff, const double* box, double* forces) {
    std::vector<char> active(ntypes, 0);
    for (int t = 0; t < ntypes; ++t)
        for (int u = 0; u < ntypes; ++u)
            if (c6[t + ntypes * u] != 0.0) active[t] = 1;
    double energy = 0.0;
    for (int i = 0; i < n; ++i) {
        if (!active[type[i]]) continue;
        for (int j = i + 1; j < n; ++j) {
            if (!active[type[j]] || mol[i] == mol[j]) continue;
            double d[3];
            for (int k = 0; k < 3; ++k) {
                d[k] = pos[3 * i + k] - pos[3 * j + k];
                if (method == 1) d[k] = periodic_diff(pos[3 * i + k], pos[3 * j + k], box[k]);
            }
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (method != 0 && r2 >= cutoff * cutoff) continue;
            const double r = std::sqrt(r2);
            const double C = c6[type[i] + ntypes * type[j]], D = d6[type[i] + ntypes * type[j]];
            const double x = D * r;
            // nested form of the xml expression (:84-87)
            const double s = 1 + (1 + (1 + (1 + (1 + (1 + (1) * x / 6) * x / 5) * x / 4) * x / 3) * x / 2) * x / 1;
            const double ex = std::exp(-x);
            const double tt6 = 1.0 - s * ex;
            const double r6 = r2 * r2 * r2;
            energy += -C * tt6 / r6;
            // dE/dr = 6 C tt6 / r^7 - C d6 x^6 e^-x / (720 r^6)
            const double x6 = x * x * x * x * x * x;
            const double dEdr = 6.0 * C * tt6 / (r6 * r) - C * D * x6 * ex / (720.0 * r6);
            for (int k = 0; k < 3; ++k) {
                const double fk = -dEdr * d[k] / r;
                forces[3 * i + k] += fk;
                forces[3 * j + k] -= fk;
            }
        }
    }
    return energy;
}

}  // namespace oracle
