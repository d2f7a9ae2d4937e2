#!/usr/bin/env python3
"""Generate the device code that evaluates the 2-body and 3-body PIPs.

E(x) = sum_k a[k] P_k(x) is a polynomial of degree <= 4 whose monomials are
listed in ``params/pip_monomials.npz``.  The monomials (variable indices
sorted, v0 <= v1 <= v2 <= v3) are arranged in a prefix tree and evaluated in
nested Horner form

    E = sum_v0 x[v0] (c_v0 + sum_v1 x[v1] (c_v0v1 + sum_v2 x[v2] (c_v0v1v2
            + sum_v3 c_v0v1v2v3 x[v3])))

with the gradient accumulated in the same sweep by the product rule: the
derivative with respect to the factor at depth d of a chain is
(product of the factors above it) * (Horner sum below it).  Every leaf costs
two fused multiply-adds (one for E, one for dE/dx[v3]), every interior node
two more plus one multiply for the running prefix product.

All 36 (3B) or 31 (2B) variables and their gradients stay in registers; the
coefficients are compile-time immediates, so one lane evaluates one dimer or
trimer with no memory traffic at all.  Each second-level block starts with an
empty-asm fence on the variables it reads, which keeps the compiler from
interleaving independent blocks (that interleaving multiplies the live
register count by the block count).  The generated function is a template on
the arithmetic type, instantiated for float (mixed-precision path) and double
(FP64 path).

The 3-body polynomial is invariant under the 24 relabellings of a trimer
that permute its waters and swap the hydrogens of an even number of them.
pip_sym_gen.h writes it as E(x) = R(x) + sum_h Q(y_h): Q holds one
representative per full orbit (1288 monomials), R the monomials of smaller
orbits (2613), so the code evaluated 24 times is small enough to stay in the
instruction cache.

Usage: gen_pip.py --out DIR  (writes DIR/pip_eval_gen.h and DIR/pip_sym_gen.h)
"""
import argparse
import json
import os
from collections import defaultdict

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PARAMS = os.path.join(HERE, '..', 'params')


def build_tree(rows, coef):
    """rows: (n, 5) int array {k, v0..v3} (-1 padded).  Returns nested dicts."""
    tree = {}
    for r in rows:
        k = int(r[0])
        vs = [int(v) for v in r[1:] if v >= 0]
        node = tree
        for d, v in enumerate(vs):
            child = node.setdefault(v, {'c': 0.0, 'kids': {}})
            if d == len(vs) - 1:
                child['c'] += coef[k]
            node = child['kids']
    return tree


def lit(v):
    r = repr(float(v))
    if 'e' not in r and '.' not in r:
        r += '.0'
    return 'T(%s)' % r


def emit(name, nvar, tree):
    """Straight-line Horner code.  Each accumulator starts from its first
    product and adds its own constant last, so no instruction needs two
    literal operands (two literals would force the compiler to keep the
    coefficient in a register, and identical coefficients -- every one recurs
    about 29 times by symmetry -- would then be merged into long-lived
    registers)."""
    out = []
    w = out.append
    stats = defaultdict(int)

    def acc(var, first, a, b):
        if first:
            w('%s = %s * %s;' % (var, a, b))
            stats['mul'] += 1
        else:
            w('%s = fma(%s, %s, %s);' % (var, a, b, var))
            stats['fma'] += 1

    def add_const(var, c):
        if c != 0.0:
            w('%s += %s;' % (var, lit(c)))
            stats['add'] += 1

    w('template <typename T>')
    w('__device__ __forceinline__ T %s(T (&x)[%d], T (&g)[%d]) {' % (name, nvar, nvar))
    w('#pragma clang fp contract(fast)')
    w('for (int i = 0; i < %d; ++i) g[i] = T(0);' % nvar)
    w('T e, s1, s2, s3, p2, p3;')
    first_e = True
    for v0 in sorted(tree):
        n1 = tree[v0]
        first1 = True
        for v1 in sorted(n1['kids']):
            n2 = n1['kids'][v1]
            # Re-define the variables this block reads (an empty asm, no
            # instruction): the Horner sums below depend only on x, and without
            # this fence the instruction selector hoists hundreds of them to the
            # top of the function, where they would all stay live at once.
            used = sorted(set([v0, v1] + list(n2['kids']) +
                              [v3 for n3 in n2['kids'].values() for v3 in n3['kids']]))
            w(' '.join('asm volatile("" : "+v"(x[%d]));' % v for v in used))
            stats['fence'] += 1
            acc('p2', True, 'x[%d]' % v0, 'x[%d]' % v1)
            first2 = True
            for v2 in sorted(n2['kids']):
                n3 = n2['kids'][v2]
                if n3['kids']:
                    acc('p3', True, 'p2', 'x[%d]' % v2)
                    first3 = True
                    for v3 in sorted(n3['kids']):
                        c = n3['kids'][v3]['c']
                        acc('s3', first3, lit(c), 'x[%d]' % v3)
                        first3 = False
                        acc('g[%d]' % v3, False, lit(c), 'p3')
                        stats['leaf'] += 1
                    add_const('s3', n3['c'])
                    acc('s2', first2, 'x[%d]' % v2, 's3')
                    acc('g[%d]' % v2, False, 'p2', 's3')
                else:
                    acc('s2', first2, lit(n3['c']), 'x[%d]' % v2)
                    acc('g[%d]' % v2, False, lit(n3['c']), 'p2')
                first2 = False
            if first2:  # level-2 node with no children: s2 is just its constant
                w('s2 = %s;' % lit(n2['c']))
            else:
                add_const('s2', n2['c'])
            acc('s1', first1, 'x[%d]' % v1, 's2')
            first1 = False
            acc('g[%d]' % v1, False, 'x[%d]' % v0, 's2')
        if first1:
            w('s1 = %s;' % lit(n1['c']))
        else:
            add_const('s1', n1['c'])
        acc('e', first_e, 'x[%d]' % v0, 's1')
        first_e = False
        w('g[%d] += s1;' % v0)
        stats['add'] += 1
    w('return e;')
    w('}')
    return '\n'.join(out), dict(stats)


# Points of the 3-body variables (reference order, kernels_pip.hip:var3):
# 0 Oa 1 Ha1 2 Ha2 3 Ob 4 Hb1 5 Hb2 6 Oc 7 Hc1 8 Hc2.
VARS3 = [(1, 2), (4, 5), (7, 8), (0, 1), (0, 2), (3, 4), (3, 5), (6, 7), (6, 8), (1, 4), (1, 5), (1, 7), (1, 8),
         (2, 4), (2, 5), (2, 7), (2, 8), (4, 7), (4, 8), (5, 7), (5, 8), (0, 4), (0, 5), (0, 7), (0, 8), (3, 1),
         (3, 2), (3, 7), (3, 8), (6, 1), (6, 2), (6, 4), (6, 5), (0, 3), (0, 6), (3, 6)]


def trimer_group():
    """Permutations of the 36 variables induced by the 24 relabellings of a
    trimer that permute the waters and swap the hydrogens of an even number
    of them: h[m] = the variable that relabelling h puts in place of m."""
    idx = {frozenset(p): n for n, p in enumerate(VARS3)}
    out = set()
    for wp in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        for sw in ((0, 0, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1)):
            pi = [0] * 9
            for w in range(3):
                t = wp[w]
                pi[3 * w] = 3 * t
                pi[3 * w + 1], pi[3 * w + 2] = (3 * t + 2, 3 * t + 1) if sw[w] else (3 * t + 1, 3 * t + 2)
            out.add(tuple(idx[frozenset((pi[a], pi[b]))] for a, b in VARS3))
    return sorted(out)


def split_by_orbits(rows, group):
    """The polynomial as E(x) = R(x) + sum_h Q(y_h), y_h[m] = x[h[m]]: Q holds one
    representative (the smallest) of every orbit of full size |group| with its
    coefficient; R holds the monomials of the smaller orbits.  Checks that every
    coefficient's monomials are closed under the group."""
    mono = {}
    for r in rows:
        m = tuple(sorted(int(v) for v in r[1:] if v >= 0))
        assert m not in mono, 'repeated monomial'
        mono[m] = int(r[0])
    orbit_rows, rest_rows, seen = [], [], set()
    for m in sorted(mono):
        if m in seen:
            continue
        orb = {tuple(sorted(h[v] for v in m)) for h in group}
        assert all(mono.get(o) == mono[m] for o in orb), 'polynomial not invariant under the group'
        seen |= orb
        pad = lambda t: list(t) + [-1] * (4 - len(t))
        if len(orb) == len(group):
            orbit_rows.append([mono[m]] + pad(min(orb)))
        else:
            rest_rows.extend([mono[o]] + pad(o) for o in sorted(orb))
    return np.array(orbit_rows), np.array(rest_rows)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--out', required=True)
    args = ap.parse_args()
    with open(os.path.join(PARAMS, 'mbpol_tables.json')) as fh:
        t = json.load(fh)
    data = np.load(os.path.join(PARAMS, 'pip_monomials.npz'))
    parts = ['// Generated by gen_pip.py from pip_monomials.npz -- do not edit.', '#pragma once', '']
    summary = {}
    for name, key, nvar, coefs in (('pip2b_eval', 'poly2b', 31, t['twobody']['thefit']),
                                   ('pip3b_eval', 'poly3b', 36, t['threebody']['thefit'])):
        code, stats = emit(name, nvar, build_tree(data[key], coefs))
        parts.append('// %s: %s' % (name, stats))
        parts.append(code)
        parts.append('')
        summary[name] = stats
    os.makedirs(args.out, exist_ok=True)
    text = '\n'.join(parts) + '\n'
    path = os.path.join(args.out, 'pip_eval_gen.h')
    if not (os.path.exists(path) and open(path).read() == text):
        with open(path, 'w') as fh:
            fh.write(text)
    # 3-body polynomial split by the trimer's relabelling group (pip_sym_gen.h)
    group = trimer_group()
    orows, rrows = split_by_orbits(data['poly3b'], group)
    coefs = t['threebody']['thefit']
    parts = ['// Generated by gen_pip.py from pip_monomials.npz -- do not edit.', '#pragma once', '',
             '// E(x) = pip3b_rest(x) + sum over the %d relabellings h of pip3b_orbit(y_h), y_h[m] = x[pip3b_perm[h][m]]'
             % len(group)]
    for name, rows in (('pip3b_orbit', orows), ('pip3b_rest', rrows)):
        code, stats = emit(name, 36, build_tree(rows, coefs))
        parts += ['// %s: %s' % (name, stats), code, '']
        summary[name] = stats
    parts.append('constexpr int pip3b_ngroup = %d;' % len(group))
    parts.append('__constant__ int pip3b_perm[%d][36] = {' % len(group))
    parts += ['    {%s},' % ', '.join(str(v) for v in h) for h in group]
    parts.append('};')
    text = '\n'.join(parts) + '\n'
    path = os.path.join(args.out, 'pip_sym_gen.h')
    if not (os.path.exists(path) and open(path).read() == text):
        with open(path, 'w') as fh:
            fh.write(text)
    print(json.dumps(summary))


if __name__ == '__main__':
    main()
