// extern "C" entry points around the reference's standalone polynomial and
// incomplete-gamma sources, so tests can call them through ctypes.  Compiled
// together with those sources (read in place from the reference tree) into
// oracle/_ref/libref_poly.so by oracle/ref/Makefile.  Test infrastructure only.
#include "poly-2b-v6x.h"
#include "poly-3b-v2x.h"
#include "gammq.h"

extern "C" {
double ref_poly_2b(const double* a, const double* x, double* g) { return poly_2b_v6x_eval(a, x, g); }
double ref_poly_3b(const double* a, const double* x, double* g) { return poly_3b_v2x::eval(a, x, g); }
double ref_gammq(double a, double x) { return ttm::gammq(a, x); }
double ref_gammln(double x) { return ttm::gammln(x); }
}
