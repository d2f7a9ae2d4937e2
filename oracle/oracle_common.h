// FP64 CPU oracle for the MB-pol energy+force path -- TEST INFRASTRUCTURE ONLY.
//
// This directory restates, in plain C++, the reference's single-threaded FP64
// "Reference" platform (platforms/reference/src/*) so that the HIP product can
// be checked against it.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it.  Each function names the reference file:line it
// follows.  Parity pins for every function are listed in oracle/README.md.
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

namespace oracle {

struct V3 {
    double x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator*(double s) const { return V3(x * s, y * s, z * s); }
    V3& operator+=(const V3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    V3& operator-=(const V3& o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
    V3& operator*=(double s) { x *= s; y *= s; z *= s; return *this; }
    double dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
    V3 cross(const V3& o) const { return V3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x); }
};

inline V3 load3(const double* p, int i) { return V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
inline void add3(double* f, int i, const V3& v) { f[3 * i] += v.x; f[3 * i + 1] += v.y; f[3 * i + 2] += v.z; }

// Minimum-image of `p` relative to `ref`, component by component
// (platforms/reference/src/MBPolReferenceTwoBodyForce.cpp:66-76).
inline void image_particle(const V3& box, const V3& ref, V3& p) {
    for (int i = 0; i < 3; ++i) {
        double d = ref[i] - p[i];
        double f = std::floor(d / box[i] + 0.5);
        p[i] += box[i] * f;
    }
}

// Molecule imaging: hydrogens relative to their own O, later O's relative to
// the first O (MBPolReferenceTwoBodyForce.cpp:78-109).  `pos` holds 3, 6 or 9
// atoms in O,H1,H2 order per water.
inline void image_molecules(const V3& box, V3* pos, int nwat) {
    image_particle(box, pos[0], pos[1]);
    image_particle(box, pos[0], pos[2]);
    for (int w = 1; w < nwat; ++w) {
        image_particle(box, pos[0], pos[3 * w]);
        image_particle(box, pos[3 * w], pos[3 * w + 1]);
        image_particle(box, pos[3 * w], pos[3 * w + 2]);
    }
}

// PIP evaluation from the monomial tables (pip.cpp).
double pip2b_eval(const double x[31], double g[31]);
double pip3b_eval(const double x[36], double g[36]);

}  // namespace oracle
