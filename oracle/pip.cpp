// FP64 evaluation of the two permutationally invariant polynomials from their
// monomial tables -- oracle (test infrastructure only).
//
// The reference evaluates E = sum_k a[k] P_k(x) and dE/dx with generated
// straight-line code (platforms/reference/src/poly-2b-v6x.cpp:3-13841,
// poly-3b-v2x.cpp:3-28378).  pip_tables.h lists every monomial of every P_k
// (one row {k, v0, v1, v2, v3} per monomial, -1 padding); here the sum is
// taken term by term in the plainest possible way.
#include "oracle_common.h"
#include "pip_tables.h"

namespace oracle {

template <int NV, int NR>
static double eval_rows(const short (&rows)[NR][5], const double* coef, const double* x, double* g) {
    for (int i = 0; i < NV; ++i) g[i] = 0.0;
    double e = 0.0;
    for (int r = 0; r < NR; ++r) {
        const double c = coef[rows[r][0]];
        int v[4];
        int deg = 0;
        for (int q = 0; q < 4; ++q)
            if (rows[r][1 + q] >= 0) v[deg++] = rows[r][1 + q];
        double prod = c;
        for (int q = 0; q < deg; ++q) prod *= x[v[q]];
        e += prod;
        // d/dx_v of c * x_v0 x_v1 ... : product of the other factors
        for (int q = 0; q < deg; ++q) {
            double p = c;
            for (int s = 0; s < deg; ++s)
                if (s != q) p *= x[v[s]];
            g[v[q]] += p;
        }
    }
    return e;
}

double pip2b_eval(const double x[31], double g[31]) {
    return eval_rows<31>(mbpol_pip::poly2b_rows, mbpol_pip::poly2b_coef, x, g);
}

double pip3b_eval(const double x[36], double g[36]) {
    return eval_rows<36>(mbpol_pip::poly3b_rows, mbpol_pip::poly3b_coef, x, g);
}

}  // namespace oracle
