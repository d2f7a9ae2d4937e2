#!/usr/bin/env python3
"""Generate the device code of the two per-water polynomials.

Both surfaces of a single water are polynomials in the same three internal
coordinates, x1 = (d1 - r_e)/r_e, x2 = (d2 - r_e)/r_e (the OH bond lengths
d1, d2 in Angstrom) and x3 = cos(theta) - cos(theta_e):

* the 1-body surface's polynomial part (Partridge-Schwenke form),
      S(x1, x2, x3) = sum_j c_j (x1^a_j x2^b_j + x1^b_j x2^a_j) x3^c_j,
  symmetric in x1 <-> x2, over the 244 non-constant entries of the
  coefficient table (the constant entry sits outside the exponential damping
  and stays with the kernel);
* the dipole-moment surface's polynomial part,
      T(x1, x2, x3) = sum_j d_j x1^a_j x2^b_j x3^c_j,
  over its 83 non-constant entries; the charge of H1 uses T(x1, x2, x3), the
  charge of H2 uses T(x2, x1, x3).

The exponent tables in ``params/mbpol_tables.json`` are 1-based (exponent + 1).
Each polynomial is expanded into its monomials, grouped by the power of x1,
then of x2, then of x3, and emitted as nested Horner evaluation with the three
partial derivatives carried through the same recurrence (for
p(x) = sum_k a_k x^k: p' <- p' x + p, then p <- p x + a_k, from the highest
power down).  Coefficients are literals; everything stays in registers.

Usage: gen_water_poly.py --out DIR   (writes DIR/water_poly_gen.h)
"""
import argparse
import json
import os
from collections import defaultdict

HERE = os.path.dirname(os.path.abspath(__file__))
PARAMS = os.path.join(HERE, '..', 'params')


def onebody_monomials(t):
    """Monomials {(i, j, k): coefficient} of S; the coefficient of entry j is the
    weighted combination of the four fitted sets (as params/gen_headers.py)."""
    ob = t['onebody']
    c = [ob['f5z'] * a + ob['fbasis'] * b + ob['fcore'] * cc + ob['frest'] * d
         for a, b, cc, d in zip(ob['c5zA'], ob['cbasis'], ob['ccore'], ob['crest'])]
    mono = defaultdict(float)
    for j in range(1, len(c)):
        a, b, k = ob['idx1'][j] - 1, ob['idx2'][j] - 1, ob['idx3'][j] - 1
        mono[(a, b, k)] += c[j]
        mono[(b, a, k)] += c[j]
    return dict(mono)


def dipole_monomials(t):
    dm = t['dipole_surface']
    mono = defaultdict(float)
    for j in range(1, len(dm['coefD'])):
        mono[(dm['idxD0'][j] - 1, dm['idxD1'][j] - 1, dm['idxD2'][j] - 1)] += dm['coefD'][j]
    return dict(mono)


def lit(v):
    r = repr(float(v))
    if 'e' not in r and '.' not in r and 'n' not in r:
        r += '.0'
    return r


class Emitter:
    def __init__(self):
        self.lines = []
        self.n = 0

    def tmp(self):
        self.n += 1
        return 't%d' % self.n

    def horner(self, mono, xs):
        """Emit code for sum_m c_m prod_d xs[d]^m[d] with all partial
        derivatives.  Returns (value expression, [derivative expressions])."""
        if not xs:
            return lit(sum(mono.values())), []
        x = xs[0]
        groups = defaultdict(dict)
        for m, cf in mono.items():
            groups[m[0]][m[1:]] = cf
        top = max(groups)
        nrest = len(xs) - 1
        v = dv = None
        drest = [None] * nrest
        for e in range(top, -1, -1):
            if e in groups:
                sv, sd = self.horner(groups[e], xs[1:])
            else:
                sv, sd = None, [None] * nrest
            if v is None:  # highest power: start from the sub-polynomial itself
                v = self.bind(sv)
                drest = [self.bind(d) for d in sd]
                continue
            # derivative in x first (it reads the old value), then the value
            dv = self.bind(v if dv is None else 'fma(%s, %s, %s)' % (dv, x, v))
            v = self.bind('%s * %s' % (v, x) if sv is None else 'fma(%s, %s, %s)' % (v, x, sv))
            for d in range(nrest):
                if drest[d] is None and sd[d] is None:
                    continue
                if drest[d] is None:
                    drest[d] = self.bind(sd[d])
                elif sd[d] is None:
                    drest[d] = self.bind('%s * %s' % (drest[d], x))
                else:
                    drest[d] = self.bind('fma(%s, %s, %s)' % (drest[d], x, sd[d]))
        return v, [dv] + drest

    def bind(self, expr):
        if expr is None:
            return None
        t = self.tmp()
        self.lines.append('    const double %s = %s;' % (t, expr))
        return t


def emit_function(name, mono, doc):
    em = Emitter()
    v, d = em.horner(mono, ['x1', 'x2', 'x3'])
    out = ['// %s' % doc,
           '// %d monomials, highest powers %s' % (len(mono), tuple(max(m[k] for m in mono) for k in range(3))),
           '__device__ __forceinline__ double %s(double x1, double x2, double x3, double& g1, double& g2, double& g3) {'
           % name]
    out += em.lines
    out.append('    g1 = %s;' % (d[0] or '0.0'))
    out.append('    g2 = %s;' % (d[1] or '0.0'))
    out.append('    g3 = %s;' % (d[2] or '0.0'))
    out.append('    return %s;' % v)
    out.append('}')
    return '\n'.join(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--out', required=True)
    args = ap.parse_args()
    with open(os.path.join(PARAMS, 'mbpol_tables.json')) as fh:
        t = json.load(fh)
    ob = t['onebody']
    c0 = ob['f5z'] * ob['c5zA'][0] + ob['fbasis'] * ob['cbasis'][0] + ob['fcore'] * ob['ccore'][0] + \
        ob['frest'] * ob['crest'][0]
    parts = ['// Generated by gen_water_poly.py from mbpol_tables.json -- do not edit.', '#pragma once', '',
             '// constant entries of the two tables (outside the exponential damping)',
             'constexpr double onebody_poly_c0 = %s;' % lit(c0),
             'constexpr double dipole_poly_d0 = %s;' % lit(t['dipole_surface']['coefD'][0]), '',
             emit_function('onebody_poly', onebody_monomials(t),
                           '1-body polynomial S(x1, x2, x3) and its gradient'), '',
             emit_function('dipole_poly', dipole_monomials(t),
                           'dipole-moment-surface polynomial T(x1, x2, x3) and its gradient'), '']
    os.makedirs(args.out, exist_ok=True)
    text = '\n'.join(parts) + '\n'
    path = os.path.join(args.out, 'water_poly_gen.h')
    if not (os.path.exists(path) and open(path).read() == text):
        with open(path, 'w') as fh:
            fh.write(text)


if __name__ == '__main__':
    main()
