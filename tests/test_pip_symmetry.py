"""The 3-body polynomial as R(x) + sum_h Q(y_h) over the 24 relabellings of a
trimer (csrc/gen_pip.py: trimer_group, split_by_orbits), the form the opt-in
HIP 3-body kernel evaluates, against the FP64 oracle's term-by-term polynomial
on random variables: energy and gradient to 1e-12 relative.  Also: the group
has 24 distinct elements and is closed under composition, every coefficient's
monomials are closed under it (split_by_orbits asserts this), and the split
keeps every monomial exactly once (|Q| * 24 + |R| = 33525)."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, '..', 'mbpol_openmm_plugin_amd', 'csrc')
PARAMS = os.path.join(HERE, '..', 'mbpol_openmm_plugin_amd', 'params')
sys.path.insert(0, CSRC)
import gen_pip  # noqa: E402


def poly(rows, coef, x):
    """sum_k coef[row k] prod x[v] and its gradient, monomial by monomial."""
    e, g = 0.0, np.zeros(len(x))
    for r in rows:
        vs = [int(v) for v in r[1:] if v >= 0]
        c = coef[int(r[0])]
        e += c * np.prod([x[v] for v in vs])
        for a in range(len(vs)):
            g[vs[a]] += c * np.prod([x[v] for b, v in enumerate(vs) if b != a])
    return e, g


def test_relabelling_split_matches_oracle(oracle):
    group = gen_pip.trimer_group()
    assert len(group) == 24 and len(set(group)) == 24
    gs = set(group)
    for h1 in group:  # closed under composition
        for h2 in group:
            assert tuple(h1[h2[m]] for m in range(36)) in gs
    data = np.load(os.path.join(PARAMS, 'pip_monomials.npz'))
    with open(os.path.join(PARAMS, 'mbpol_tables.json')) as fh:
        coef = json.load(fh)['threebody']['thefit']
    orows, rrows = gen_pip.split_by_orbits(data['poly3b'], group)
    assert len(orows) * 24 + len(rrows) == len(data['poly3b'])
    rng = np.random.default_rng(4)
    for _ in range(3):
        x = rng.uniform(0.05, 0.9, 36)
        e_ref, g_ref = oracle.pip3b(x)
        e, g = poly(rrows, coef, x)
        for h in group:
            y = x[list(h)]
            eq, gq = poly(orows, coef, y)
            e += eq
            np.add.at(g, list(h), gq)
        print('E %.15g vs oracle %.15g; gradient rel err %.1e' %
              (e, e_ref, np.abs(g - g_ref).max() / np.abs(g_ref).max()), flush=True)
        assert abs(e - e_ref) <= 1e-12 * abs(e_ref)
        assert np.abs(g - g_ref).max() <= 1e-12 * np.abs(g_ref).max()
