#!/usr/bin/env python3
"""Extract the MB-pol force-field tables from the reference tree into data files.

Run once, in a container where the reference tree is mounted (default
``/root/reference``).  The outputs are plain data committed next to this
script; nothing at build or run time reads the reference tree.

What is extracted (all of it numbers the potential is defined by, not code):

* ``mbpol_tables.json`` -- scalar constants and coefficient tables:
    - 1-body Partridge-Schwenke-type monomer surface
      (``platforms/reference/src/mbpol_interaction_constants.h:4-423``),
    - 2-body variable constants, switch radii and ``thefit[1153]``
      (``platforms/reference/src/mbpol_2body_constants.h:34-1208``),
    - 3-body variable constants, switch radii and ``thefit[1163]``
      (``platforms/reference/src/mbpol_3body_constants.h:5-1186``),
    - the dipole-moment surface used for the geometry-dependent charges
      (``platforms/reference/src/MBPolReferenceElectrostaticsForce.cpp:2796-2877``),
    - force-field parameters: charges / damping / polarisability, Thole
      factors, cutoffs, M-site weights and the dispersion C6/d6 tables
      (``python/mbpol.xml:14-83``).
* ``pip_monomials.npz`` -- the monomial structure of the two permutationally
  invariant polynomials.  The reference ships them as machine-generated,
  reverse-mode differentiated straight-line code
  (``poly-2b-v6x.cpp:3-13841``, ``poly-3b-v2x.cpp:3-28378``).  The energy is
  linear in the coefficient vector ``a``: ``E(x) = sum_k a[k] * P_k(x)``.
  This script symbolically executes the energy expression (the ``return``
  statement and the ``t`` temporaries it depends on) and records, for every
  coefficient ``k``, the set of monomials of ``P_k``.  Every monomial turns
  out to carry multiplicity 1, which the script asserts.

Usage::

    python3 extract_reference_params.py [--reference /root/reference]
"""
import argparse
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _read(ref, rel):
    with open(os.path.join(ref, rel)) as fh:
        return fh.read()


def _scalar(text, name):
    m = re.search(r'const\s+(?:double|size_t)\s+' + re.escape(name) + r'\s*=\s*([-+0-9.eE]+)\s*;', text)
    if not m:
        raise KeyError(name)
    return m.group(1)


def _array(text, name, count):
    m = re.search(r'const\s+(?:double|size_t)\s+' + re.escape(name) + r'\s*\[\s*\d*\s*\]\s*=\s*\{(.*?)\}\s*;', text, re.S)
    if not m:
        raise KeyError(name)
    body = re.sub(r'//[^\n]*', '', m.group(1))
    vals = [v for v in re.split(r'[\s,]+', body) if v]
    if len(vals) != count:
        raise ValueError('%s: expected %d values, found %d' % (name, count, len(vals)))
    return vals


def _float_list(vals):
    out = [float(v) for v in vals]
    return out


# ---------------------------------------------------------------------------
# symbolic execution of the generated polynomial code
# ---------------------------------------------------------------------------

_TOK = re.compile(r'\s*(t\d+|a\[\d+\]|x\[\d+\]|\d+\.\d*|\d+|[()+*-])')


def _tokenize(expr):
    pos, out = 0, []
    while pos < len(expr):
        m = _TOK.match(expr, pos)
        if not m:
            raise ValueError('cannot tokenize: ' + expr[pos:pos + 30])
        out.append(m.group(1))
        pos = m.end()
    return out


class _PolyTracer:
    """Polynomials are dicts {(coef_index or -1, sorted var tuple): multiplier}."""

    def __init__(self, defs):
        self.defs = defs
        self.memo = {}

    @staticmethod
    def _mul(p, q):
        r = {}
        for (k1, m1), c1 in p.items():
            for (k2, m2), c2 in q.items():
                if k1 >= 0 and k2 >= 0:
                    raise ValueError('expression is not linear in the coefficients')
                key = (max(k1, k2), tuple(sorted(m1 + m2)))
                r[key] = r.get(key, 0) + c1 * c2
        return r

    @staticmethod
    def _add(p, q, sign):
        r = dict(p)
        for key, c in q.items():
            r[key] = r.get(key, 0) + sign * c
        return r

    def value(self, name):
        if name not in self.memo:
            self.memo[name] = self.parse(self.defs[name])
        return self.memo[name]

    def parse(self, expr):
        toks = _tokenize(expr)
        pos = [0]

        def peek():
            return toks[pos[0]] if pos[0] < len(toks) else None

        def take():
            t = toks[pos[0]]
            pos[0] += 1
            return t

        def sum_():
            neg = False
            if peek() == '-':
                take()
                neg = True
            p = prod()
            if neg:
                p = {k: -v for k, v in p.items()}
            while peek() in ('+', '-'):
                op = take()
                p = self._add(p, prod(), 1 if op == '+' else -1)
            return p

        def prod():
            p = atom()
            while peek() == '*':
                take()
                p = self._mul(p, atom())
            return p

        def atom():
            t = take()
            if t == '(':
                p = sum_()
                if take() != ')':
                    raise ValueError('unbalanced parentheses')
                return p
            if t[0] == 't':
                return self.value(t)
            if t.startswith('a['):
                return {(int(t[2:-1]), ()): 1}
            if t.startswith('x['):
                return {(-1, (int(t[2:-1]),)): 1}
            return {(-1, ()): float(t)}

        p = sum_()
        if peek() is not None:
            raise ValueError('trailing tokens in expression')
        return p


def trace_polynomial(source_text, n_coef, n_var):
    body = source_text[source_text.index('{') + 1:]
    defs, ret = {}, None
    for stmt in body.split(';'):
        s = ' '.join(stmt.split())
        if s.startswith('const double t'):
            name, expr = s[len('const double '):].split('=', 1)
            defs[name.strip()] = expr.strip()
        elif s.startswith('return'):
            ret = s[len('return'):].strip()
    sys.setrecursionlimit(200000)
    energy = _PolyTracer(defs).parse(ret)
    rows = []
    for (k, mono), c in energy.items():
        if c == 0:
            continue
        if k < 0:
            raise ValueError('coefficient-free term in the energy')
        if c != 1:
            raise ValueError('monomial multiplicity %r != 1' % (c,))
        if not (1 <= len(mono) <= 4):
            raise ValueError('unexpected monomial degree %d' % len(mono))
        if max(mono) >= n_var or k >= n_coef:
            raise ValueError('index out of range')
        rows.append([k] + list(mono) + [-1] * (4 - len(mono)))
    rows.sort(key=lambda r: (r[1], r[2], r[3], r[4], r[0]))
    return np.array(rows, dtype=np.int16)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--reference', default='/root/reference')
    ap.add_argument('--out', default=HERE)
    args = ap.parse_args()
    ref = args.reference
    src = 'platforms/reference/src/'

    c1 = _read(ref, src + 'mbpol_interaction_constants.h')
    c2 = _read(ref, src + 'mbpol_2body_constants.h')
    c3 = _read(ref, src + 'mbpol_3body_constants.h')
    el = _read(ref, src + 'MBPolReferenceElectrostaticsForce.cpp')
    xml = _read(ref, 'python/mbpol.xml')

    tables = {'_comment': 'MB-pol force-field tables extracted as data by extract_reference_params.py; '
                          'see that script for the reference file:line of each table.'}

    # --- 1-body (mbpol_interaction_constants.h) ---
    one = {}
    for name in ('Eh_J', 'Na', 'kcal_J', 'Bohr_A', 'c0', 'ea0', 'h_Js', 'hbar_Js',
                 'reoh', 'thetae', 'b1', 'roh', 'alphaoh', 'deohA', 'phh1A', 'phh2'):
        one[name] = float(_scalar(c1, name))
    for name in ('c5zA', 'cbasis', 'ccore', 'crest'):
        one[name] = _float_list(_array(c1, name, 245))
    for name in ('idx1', 'idx2', 'idx3'):
        one[name] = [int(v) for v in _array(c1, name, 245)]
    # scale factors and fixed numbers that live inside calculateOneBodyIxn
    ob = _read(ref, src + 'MBPolReferenceOneBodyForce.cpp')
    for name in ('f5z', 'fbasis', 'fcore', 'frest', 'costhe'):
        one[name] = float(_scalar(ob, name))
    m = re.search(r'e1\s*\+=\s*([-+0-9.eE]+)\s*;', ob)
    one['e1_shift'] = float(m.group(1))
    tables['onebody'] = one

    # --- 2-body ---
    two = {}
    for name in ('k_HH_intra', 'k_OH_intra', 'k_HH_coul', 'k_OH_coul', 'k_OO_coul',
                 'k_XH_main', 'k_XO_main', 'k_XX_main', 'in_plane_gamma',
                 'out_of_plane_gamma', 'r2i', 'r2f'):
        two[name] = float(_scalar(c2, name))
    m = re.search(r'const double thefit\[\]\s*=\s*\{(.*?)\};', c2, re.S)
    vals = re.findall(r'([-+]?\d\.\d+e[-+]\d+)', re.sub(r'//[^\n]*', '', m.group(1)))
    assert len(vals) == 1153, len(vals)
    two['thefit'] = _float_list(vals)
    tb = _read(ref, src + 'MBPolReferenceTwoBodyForce.cpp')
    two['d0_intra'] = float(_scalar(tb, 'd0_intra'))
    two['d0_inter'] = float(_scalar(tb, 'd0_inter'))
    tables['twobody'] = two

    # --- 3-body ---
    three = {}
    for name in ('r3i', 'r3f', 'kHH_intra', 'kOH_intra', 'kHH', 'kOH', 'kOO',
                 'dHH_intra', 'dOH_intra', 'dHH', 'dOH', 'dOO'):
        three[name] = float(_scalar(c3, name))
    m = re.search(r'const double thefit\[\]\s*=\s*\{(.*?)\};', c3, re.S)
    vals = re.findall(r'([-+]?\d\.\d+e[-+]\d+)', re.sub(r'//[^\n]*', '', m.group(1)))
    assert len(vals) == 1163, len(vals)
    three['thefit'] = _float_list(vals)
    tables['threebody'] = three

    # --- dipole-moment surface (computeWaterCharge) ---
    wc_start = el.index('void MBPolReferenceElectrostaticsForce::computeWaterCharge(')
    wc = el[wc_start:]
    dms = {}
    for name in ('gammaM', 'ath0', 'costhe', 'reoh', 'b1D', 'a', 'b', 'c0', 'c1', 'c2', 'Bohr_A'):
        dms[name] = float(_scalar(wc, name))
    dms['coefD'] = _float_list(_array(wc, 'coefD', 84))
    for name in ('idxD0', 'idxD1', 'idxD2'):
        dms[name] = [int(v) for v in _array(wc, name, 84)]
    tables['dipole_surface'] = dms

    # --- electrostatics constants (MBPolReferenceElectrostaticsForce.cpp) ---
    est = {}
    m = re.search(r'_electric\(([-+0-9.eE]+)\)', el)
    est['electric'] = float(m.group(1))
    m = re.search(r'_polarSOR\(([-+0-9.eE]+)\)', el)
    est['polar_sor'] = float(m.group(1))
    m = re.search(r'_debye\(([-+0-9.eE]+)\)', el)
    est['debye'] = float(m.group(1))
    m = re.search(r'SQRT_PI\s*=\s*([-+0-9.eE]+)\s*;', el)
    est['sqrt_pi_truncated'] = float(m.group(1))
    m = re.search(r'MBPOL_PME_ORDER\s*=\s*(\d+)\s*;', el)
    est['pme_order'] = int(m.group(1))
    tables['electrostatics_constants'] = est

    # --- force-field parameters (python/mbpol.xml) ---
    ff = {}
    m = re.search(r'weight1="([^"]+)" weight2="([^"]+)" weight3="([^"]+)"', xml)
    ff['msite_weights'] = [float(m.group(i)) for i in (1, 2, 3)]
    m = re.search(r'<MBPolElectrostaticsForce thole-charge-charge="([^"]+)" thole-charge-dipole="([^"]+)" '
                  r'thole-dipole-dipole="([^"]+)" thole-dipole-dipole-singlebond="([^"]+)"', xml)
    cc, cd, dd, ddsb = (float(m.group(i)) for i in (1, 2, 3, 4))
    # order TCC, TCD, TDD, TDDOH, TDDHH as assembled by python/mbpol.py:266-269
    ff['thole'] = [cc, cd, dd, ddsb, dd]
    atoms = {}
    for t, q, d, p in re.findall(r'<Atom type="MBPol-(\w+)" charge="([^"]+)" damping-factor="([^"]+)" '
                                 r'polarizability="([^"]+)"', xml):
        atoms[t] = {'charge': float(q), 'damping': float(d), 'polarizability': float(p)}
    ff['electrostatics_atoms'] = atoms
    ff['twobody_cutoff_nm'] = float(re.search(r'<MBPolTwoBodyForce cutoff_nm="([^"]+)"', xml).group(1))
    ff['threebody_cutoff_nm'] = float(re.search(r'<MBPolThreeBodyForce cutoff_nm="([^"]+)"', xml).group(1))

    def disp_table(name):
        m = re.search(name + r'\s*=\s*\[(.*?)\]', xml, re.S)
        body = re.sub(r'#[^\n]*', '', m.group(1))
        vals = [float(v) for v in re.split(r'[\s,]+', body) if v]
        assert len(vals) == 16, (name, len(vals))
        return vals
    ff['dispersion_types'] = ['O', 'H', 'M', 'CL']
    ff['dispersion_C6'] = disp_table('C6table')
    ff['dispersion_d6'] = disp_table('d6table')
    tables['forcefield'] = ff

    with open(os.path.join(args.out, 'mbpol_tables.json'), 'w') as fh:
        json.dump(tables, fh, indent=1)

    # --- polynomial monomial structure ---
    p2 = trace_polynomial(_read(ref, src + 'poly-2b-v6x.cpp'), 1153, 31)
    p3 = trace_polynomial(_read(ref, src + 'poly-3b-v2x.cpp'), 1163, 36)
    np.savez_compressed(os.path.join(args.out, 'pip_monomials.npz'), poly2b=p2, poly3b=p3)
    print('2B monomials: %d, 3B monomials: %d' % (len(p2), len(p3)))


if __name__ == '__main__':
    main()
