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
pip_orbit_gen.h writes it as E(x) = R(x) + sum_h Q(y_h): Q holds one
representative of every orbit whose stabiliser order is a power of two (1518
of 1526 orbits), with its coefficient divided by that order, and R the 33
monomials of the other orbits.  The kernel runs Q in a loop of 24 passes,
permuting the variable and gradient registers between passes along a
Hamiltonian cycle of the group, so the code of one pass (a few thousand
instructions) stays in the instruction cache, and Q reads its coefficients
from a table in order (scalar loads) instead of carrying them as
instruction literals.

The 2-body polynomial is written the same way over the 8 relabellings of a
dimer (swap the waters; swap the hydrogens, and with them the two extra points,
of either water): 1800 orbit representatives, all with stabilisers of order
1, 2, 4 or 8, so R is empty and the kernel runs 8 passes of Q.

Usage: gen_pip.py --out DIR  (writes DIR/pip_eval_gen.h and DIR/pip_orbit_gen.h)
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


def _lit(v):
    r = repr(float(v))
    if 'e' not in r and '.' not in r:
        r += '.0'
    return 'T(%s)' % r


def emit(name, nvar, tree, table=None, accumulate=False):
    """Straight-line Horner code.  Each accumulator starts from its first
    product and adds its own constant last, so no instruction needs two
    literal operands (two literals would force the compiler to keep the
    coefficient in a register, and identical coefficients -- every one recurs
    about 29 times by symmetry -- would then be merged into long-lived
    registers).

    table: None -- coefficients are literals; a list -- every coefficient use
    appends its value to the list and reads it as cf[k] (k = its position), so
    the function takes the table as an argument and reads it in order.
    accumulate: add the gradient into g instead of overwriting it."""
    out = []
    w = out.append
    stats = defaultdict(int)

    def lit(v):
        if table is None:
            return _lit(v)
        table.append(float(v))
        return 'cf[%d]' % (len(table) - 1)

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

    if table is None:
        w('template <typename T>')
        w('__device__ __forceinline__ T %s(T (&x)[%d], T (&g)[%d]) {' % (name, nvar, nvar))
    else:  # cf: any pointer type (e.g. one into the constant address space)
        w('template <typename T, typename P>')
        w('__host__ __device__ __forceinline__ T %s(T (&x)[%d], T (&g)[%d], P cf) {' % (name, nvar, nvar))
    w('#pragma clang fp contract(fast)')
    if not accumulate:
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
            w(' '.join('PIP_FENCE(x[%d]);' % v for v in used))
            stats['fence'] += 1
            acc('p2', True, 'x[%d]' % v0, 'x[%d]' % v1)
            first2 = True
            for v2 in sorted(n2['kids']):
                n3 = n2['kids'][v2]
                if n3['kids']:
                    acc('p3', True, 'p2', 'x[%d]' % v2)
                    first3 = True
                    for v3 in sorted(n3['kids']):
                        c = lit(n3['kids'][v3]['c'])
                        acc('s3', first3, c, 'x[%d]' % v3)
                        first3 = False
                        acc('g[%d]' % v3, False, c, 'p3')
                        stats['leaf'] += 1
                    add_const('s3', n3['c'])
                    acc('s2', first2, 'x[%d]' % v2, 's3')
                    acc('g[%d]' % v2, False, 'p2', 's3')
                else:
                    c = lit(n3['c'])
                    acc('s2', first2, c, 'x[%d]' % v2)
                    acc('g[%d]' % v2, False, c, 'p2')
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
    w('return %s;' % ('T(0)' if first_e else 'e'))  # an empty polynomial is 0
    w('}')
    if table is not None:
        out = prefetch_table_reads(out, len(table))
    return '\n'.join(out), dict(stats)


PREFETCH_BATCH = 16


def prefetch_table_reads(lines, ncoef, batch=PREFETCH_BATCH):
    """Rewrite the cf[k] reads of a table-mode function as explicit loads into
    named constants, batch by batch, with batch j+1 loaded right after the
    first instruction that uses batch j.  Scalar loads of the coefficient table
    may complete out of order, so a use waits for every load in flight: issued
    after that first use, the next batch's loads are covered by the rest of the
    current batch's arithmetic instead of stalling the next use.  Scheduling
    barriers keep the compiler from moving the loads to their uses."""
    import re
    pat = re.compile(r'cf\[(\d+)\]')
    nb = (ncoef + batch - 1) // batch

    def loads(j):
        if j >= nb:
            return []
        ks = range(j * batch, min(ncoef, (j + 1) * batch))
        return ([' '.join('const T c%d = cf[%d];' % (k, k) for k in ks), '__builtin_amdgcn_sched_barrier(0);'])

    out, started = [], set()
    for ln in lines:
        ks = [int(m) for m in pat.findall(ln)]
        if not ks:
            out.append(ln)
            continue
        ln = pat.sub(lambda m: 'c' + m.group(1), ln)
        j = ks[0] // batch
        if j == 0 and 0 not in started:
            out += loads(0)
        out.append(ln)
        if j not in started:
            started.add(j)
            out += ['__builtin_amdgcn_sched_barrier(0);'] + loads(j + 1)
    return out


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


# 2-body points: 0 Oa 1 Ha1 2 Ha2 3 Ob 4 Hb1 5 Hb2 6 Xa1 7 Xa2 8 Xb1 9 Xb2 (X: the extra
# points of each water); variables in the reference order (mbpol_2body_constants.cpp)
VARS2 = [(1, 2), (4, 5), (0, 1), (0, 2), (3, 4), (3, 5), (1, 4), (1, 5), (2, 4), (2, 5), (0, 4), (0, 5), (3, 1),
         (3, 2), (0, 3), (6, 4), (6, 5), (7, 4), (7, 5), (8, 1), (8, 2), (9, 1), (9, 2), (0, 8), (0, 9), (3, 6),
         (3, 7), (6, 8), (6, 9), (7, 8), (7, 9)]


def dimer_group():
    """Permutations of the 31 variables induced by the 8 relabellings of a
    dimer: swap the waters, swap the hydrogens of either water (its two extra
    points, which sit on either side of the molecular plane, swap with them)."""
    idx = {frozenset(p): n for n, p in enumerate(VARS2)}
    out = set()
    for swap_w in (0, 1):
        for sa in (0, 1):
            for sb in (0, 1):
                wa = [0, 2, 1, 7, 6] if sa else [0, 1, 2, 6, 7]  # O, H1, H2, X1, X2 of water a
                wb = [3, 5, 4, 9, 8] if sb else [3, 4, 5, 8, 9]
                src = wb + wa if swap_w else wa + wb
                pi = [0] * 10
                for t, q in zip([0, 1, 2, 6, 7, 3, 4, 5, 8, 9], src):
                    pi[t] = q
                out.add(tuple(idx[frozenset((pi[a], pi[b]))] for a, b in VARS2))
    return sorted(out)


def compose(h, s):
    """The relabelling 'h, then s' as a variable map: (h o s)[m] = h[s[m]]."""
    return tuple(h[s[m]] for m in range(len(s)))


def group_cycle(group):
    """A Hamiltonian cycle of the group's right Cayley graph over two
    generators a, b: steps s_1..s_n in {a, b} such that the prefixes
    h_k = s_1 o ... o s_k visit every element once and h_n = identity.  A loop
    that applies s_k to the registers after its k-th pass therefore sees every
    relabelling once and ends in the original labelling.  Among all generator
    pairs with such a cycle, the one with the fewest register moves (36 minus
    fixed points plus cycles, per permuted vector) is taken.
    Returns (a, b, mask) with bit k of mask set when step k+1 is a."""
    ident = tuple(range(len(group[0])))

    def moves(p):
        seen, cyc = set(), 0
        for i in range(len(p)):
            if i in seen or p[i] == i:
                continue
            cyc += 1
            while i not in seen:
                seen.add(i)
                i = p[i]
        return sum(1 for i in range(len(p)) if p[i] != i) + cyc

    def hamilton(a, b):
        path, used, steps = [ident], {ident}, []

        def dfs():
            if len(path) == len(group):
                for st, bit in ((a, 1), (b, 0)):
                    if compose(path[-1], st) == ident:
                        steps.append(bit)
                        return True
                return False
            for st, bit in ((a, 1), (b, 0)):
                nx = compose(path[-1], st)
                if nx in used:
                    continue
                used.add(nx)
                path.append(nx)
                steps.append(bit)
                if dfs():
                    return True
                used.discard(nx)
                path.pop()
                steps.pop()
            return False
        return steps if dfs() else None

    best = None
    for a in group:
        for b in group:
            if ident in (a, b) or a >= b:
                continue
            steps = hamilton(a, b)
            if steps is None:
                continue
            cost = sum(moves(a) if bit else moves(b) for bit in steps)
            if best is None or cost < best[0]:
                best = (cost, a, b, steps)
    _, a, b, steps = best
    return a, b, sum(bit << k for k, bit in enumerate(steps))


def orbit_loop_split(rows, coefs, group):
    """The polynomial as E(x) = R(x) + sum_h Q(y_h), y_h[m] = x[h[m]].

    Summing a monomial m over all relabellings gives |Stab(m)| times the sum of
    its orbit, so every orbit is represented in Q by its smallest member with
    coefficient c/|Stab(m)|.  That division is exact for the stabiliser orders
    1, 2, 4, 8; the few orbits with stabilisers of order 3, 6 or 24 are written
    out monomial by monomial in R with their own coefficients.  Checks that
    every coefficient's monomials are closed under the group.  Returns
    (Q rows, Q coefficients, R rows, R coefficients); rows are {k, v0..v3} with
    k indexing the returned coefficient list."""
    mono = {}
    for r in rows:
        m = tuple(sorted(int(v) for v in r[1:] if v >= 0))
        assert m not in mono, 'repeated monomial'
        mono[m] = int(r[0])
    pad = lambda t: list(t) + [-1] * (4 - len(t))
    qrows, qcoef, rrows, rcoef, seen = [], [], [], [], set()
    for m in sorted(mono):
        if m in seen:
            continue
        orb = {tuple(sorted(h[v] for v in m)) for h in group}
        assert all(mono.get(o) == mono[m] for o in orb), 'polynomial not invariant under the group'
        seen |= orb
        stab = len(group) // len(orb)
        c = coefs[mono[m]]
        if stab & (stab - 1) == 0:
            qrows.append([len(qcoef)] + pad(min(orb)))
            qcoef.append(c / stab)
        else:
            for o in sorted(orb):
                rrows.append([len(rcoef)] + pad(o))
                rcoef.append(c)
    return np.array(qrows), qcoef, np.array(rrows), rcoef


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--out', required=True)
    args = ap.parse_args()
    with open(os.path.join(PARAMS, 'mbpol_tables.json')) as fh:
        t = json.load(fh)
    data = np.load(os.path.join(PARAMS, 'pip_monomials.npz'))
    parts = ['// Generated by gen_pip.py from pip_monomials.npz -- do not edit.', '#pragma once', '',
             '// PIP_FENCE(v): an empty asm that re-defines v in a register (the includer defines it)', '']
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
    # the 2- and 3-body polynomials as loops over their relabelling groups (pip_orbit_gen.h)
    parts = ['// Generated by gen_pip.py from pip_monomials.npz -- do not edit.', '#pragma once', '']
    for prefix, key, nvar, coefs, group in (('pip2b', 'poly2b', 31, t['twobody']['thefit'], dimer_group()),
                                            ('pip3b', 'poly3b', 36, t['threebody']['thefit'], trimer_group())):
        qrows, qcoef, rrows, rcoef = orbit_loop_split(data[key], coefs, group)
        a, b, mask = group_cycle(group)
        table = []
        qcode, qstats = emit(prefix + '_orbit_pass', nvar, build_tree(qrows, qcoef), table=table, accumulate=True)
        rcode, rstats = emit(prefix + '_rest', nvar, build_tree(rrows, rcoef))
        summary[prefix + '_orbit_pass'] = qstats
        summary[prefix + '_rest'] = rstats
        parts += ['// E(x) = %s_rest(x) + sum over the %d relabellings h of %s_orbit_pass(y_h), y_h[m] = x[h[m]].'
                  % (prefix, len(group), prefix),
                  '// The relabellings are visited by applying %s_step_a or %s_step_b to the registers after'
                  % (prefix, prefix),
                  '// each pass (pass k is followed by step a when bit k of %s_step_mask is set); the last step'
                  % prefix,
                  '// returns to the original labelling.  %s_orbit_pass reads its %d coefficients in order from'
                  % (prefix, len(table)),
                  '// %s_orbit_coef (double or float table) and adds its gradient into g.' % prefix,
                  'constexpr int %s_npass = %d;' % (prefix, len(group)),
                  'constexpr unsigned %s_step_mask = 0x%xu;' % (prefix, mask),
                  'constexpr int %s_step_a[%d] = {%s};' % (prefix, nvar, ', '.join(str(v) for v in a)),
                  'constexpr int %s_step_b[%d] = {%s};' % (prefix, nvar, ', '.join(str(v) for v in b)),
                  'constexpr int %s_orbit_ncoef = %d;' % (prefix, len(table))]
        for typ, suf in (('double', 'f64'), ('float', 'f32')):
            parts.append('__constant__ %s %s_orbit_coef_%s[%d] = {' % (typ, prefix, suf, len(table)))
            for i in range(0, len(table), 4):
                parts.append('    ' + ', '.join(repr(float(v)) for v in table[i:i + 4]) + ',')
            parts.append('};')
        parts += ['', '// %s' % qstats, qcode, '', '// %s' % rstats,
                  rcode.replace('__device__', '__host__ __device__'), '']
    text = '\n'.join(parts) + '\n'
    path = os.path.join(args.out, 'pip_orbit_gen.h')
    if not (os.path.exists(path) and open(path).read() == text):
        with open(path, 'w') as fh:
            fh.write(text)
    print(json.dumps(summary))


if __name__ == '__main__':
    main()
