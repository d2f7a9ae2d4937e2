"""The product's per-lane device math compiled for the host (tests/host/
host_kernels.cpp, g++ on the same headers the HIP kernels include) against
the FP64 oracle and the reference's golden polynomial vectors.  CPU only.

* The 3-body polynomial as the kernel evaluates it -- a loop of 24 passes of
  one orbit-representative polynomial over the trimer's relabelling group,
  with the registers permuted between passes (csrc/pip_orbit_loop.h) -- must
  give the reference's energy and gradient (tests/golden/pip_vectors.json,
  written from the reference's own poly-3b-v2x.cpp) to 1e-11, and the
  oracle's to 1e-12 on random variables.  The relabelling walk must visit
  every group element once and return to the identity, and the split must
  account for every monomial of the polynomial.
* The same for the 2-body polynomial as its kernel evaluates it: 8 passes of
  one orbit-representative polynomial over the dimer's relabelling group
  (swap the waters; swap the hydrogens, and with them the extra points, of
  either water), against tests/golden/pip_vectors.json (written from the
  reference's poly-2b-v6x.cpp) and the oracle.
* The per-water terms (csrc/water_terms.h: 1-body energy and gradient,
  dipole-moment-surface charges and their derivatives) against the oracle's
  restatements of MBPolReferenceOneBodyForce.cpp:69-201 and
  MBPolReferenceElectrostaticsForce.cpp:2793-3096 on distorted waters, to
  1e-12 relative; the oracle is pinned to the reference's known answers in
  tests/test_oracle_reference_kats.py.
"""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, '..'))
CSRC = os.path.join(ROOT, 'mbpol_openmm_plugin_amd', 'csrc')
PARAMS = os.path.join(ROOT, 'mbpol_openmm_plugin_amd', 'params')
GOLDEN = os.path.join(HERE, 'golden')


@pytest.fixture(scope='module')
def host(tmp_path_factory):
    d = str(tmp_path_factory.mktemp('host_kernels'))
    gen = os.path.join(d, 'gen')
    for cmd in ([sys.executable, os.path.join(PARAMS, 'gen_headers.py'), '--out', gen],
                [sys.executable, os.path.join(CSRC, 'gen_pip.py'), '--out', gen],
                [sys.executable, os.path.join(CSRC, 'gen_water_poly.py'), '--out', gen]):
        subprocess.check_call(cmd, stdout=subprocess.DEVNULL)
    lib = os.path.join(d, 'libhost_kernels.so')
    subprocess.check_call(['g++', '-O2', '-std=c++17', '-fPIC', '-shared', '-ffp-contract=off', '-Wno-unknown-pragmas',
                           '-I' + gen, '-I' + CSRC, os.path.join(HERE, 'host', 'host_kernels.cpp'), '-o', lib])
    L = ctypes.CDLL(lib)
    dp = ctypes.POINTER(ctypes.c_double)
    L.host_pip3b_orbit.restype = ctypes.c_double
    L.host_pip3b_orbit.argtypes = [dp, dp]
    L.host_pip2b_orbit.restype = ctypes.c_double
    L.host_pip2b_orbit.argtypes = [dp, dp]
    L.host_water_onebody.restype = ctypes.c_double
    L.host_water_onebody.argtypes = [dp, dp]
    L.host_water_charges.argtypes = [dp, dp, dp]
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def orbit_eval(host, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    g = np.zeros(36)
    return host.host_pip3b_orbit(_p(x), _p(g)), g


def test_orbit_loop_matches_reference_vectors(host):
    d = json.load(open(os.path.join(GOLDEN, 'pip_vectors.json')))
    for v in d['poly3b']:
        e, g = orbit_eval(host, np.array(v['x']))
        ref = np.array(v['g'])
        assert abs(e - v['e']) <= 1e-11 * max(1.0, abs(v['e']))
        np.testing.assert_allclose(g, ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())


def test_orbit_loop_matches_oracle(host, oracle):
    rng = np.random.default_rng(11)
    for _ in range(8):
        x = rng.uniform(0.02, 1.2, 36)
        e, g = orbit_eval(host, x)
        e_ref, g_ref = oracle.pip3b(x)
        print('E %.15g vs %.15g, grad rel err %.1e' % (e, e_ref, np.abs(g - g_ref).max() / np.abs(g_ref).max()))
        assert abs(e - e_ref) <= 1e-12 * abs(e_ref)
        assert np.abs(g - g_ref).max() <= 1e-12 * np.abs(g_ref).max()
    # a single variable perturbed: the loop must not mix up which register holds which variable
    x = rng.uniform(0.02, 1.2, 36)
    for m in (0, 9, 21, 33, 35):
        y = x.copy()
        y[m] *= 1.37
        e, g = orbit_eval(host, y)
        e_ref, g_ref = oracle.pip3b(y)
        assert abs(e - e_ref) <= 1e-12 * abs(e_ref)
        assert np.abs(g - g_ref).max() <= 1e-12 * np.abs(g_ref).max()


def orbit2_eval(host, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    g = np.zeros(31)
    return host.host_pip2b_orbit(_p(x), _p(g)), g


def test_dimer_orbit_loop_matches_reference_vectors(host):
    d = json.load(open(os.path.join(GOLDEN, 'pip_vectors.json')))
    for v in d['poly2b']:
        e, g = orbit2_eval(host, np.array(v['x']))
        ref = np.array(v['g'])
        assert abs(e - v['e']) <= 1e-11 * max(1.0, abs(v['e']))
        np.testing.assert_allclose(g, ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())


def test_dimer_orbit_loop_matches_oracle(host, oracle):
    rng = np.random.default_rng(12)
    for _ in range(8):
        x = rng.uniform(0.02, 1.2, 31)
        e, g = orbit2_eval(host, x)
        e_ref, g_ref = oracle.pip2b(x)
        print('E %.15g vs %.15g, grad rel err %.1e' % (e, e_ref, np.abs(g - g_ref).max() / np.abs(g_ref).max()))
        assert abs(e - e_ref) <= 1e-12 * abs(e_ref)
        assert np.abs(g - g_ref).max() <= 1e-12 * np.abs(g_ref).max()
    x = rng.uniform(0.02, 1.2, 31)
    for m in (0, 6, 14, 15, 27, 30):
        y = x.copy()
        y[m] *= 1.37
        e, g = orbit2_eval(host, y)
        e_ref, g_ref = oracle.pip2b(y)
        assert abs(e - e_ref) <= 1e-12 * abs(e_ref)
        assert np.abs(g - g_ref).max() <= 1e-12 * np.abs(g_ref).max()


def test_dimer_orbit_split_and_walk():
    sys.path.insert(0, CSRC)
    import gen_pip
    group = gen_pip.dimer_group()
    assert len(group) == 8 and len(set(group)) == 8
    gs = set(group)
    for h1 in group:
        for h2 in group:
            assert gen_pip.compose(h1, h2) in gs
    a, b, mask = gen_pip.group_cycle(group)
    h, visited = tuple(range(31)), []
    for k in range(8):
        visited.append(h)
        h = gen_pip.compose(h, a if (mask >> k) & 1 else b)
    assert h == tuple(range(31)) and sorted(visited) == sorted(group)
    data = np.load(os.path.join(PARAMS, 'pip_monomials.npz'))
    with open(os.path.join(PARAMS, 'mbpol_tables.json')) as fh:
        coef = json.load(fh)['twobody']['thefit']
    qrows, qcoef, rrows, rcoef = gen_pip.orbit_loop_split(data['poly2b'], coef, group)
    assert len(rrows) == 0  # every stabiliser order is a power of two
    covered = {}
    for r in qrows:
        m = tuple(int(v) for v in r[1:] if v >= 0)
        for o in {tuple(sorted(hh[v] for v in m)) for hh in group}:
            covered[o] = covered.get(o, 0) + 1
    assert len(covered) == len(data['poly2b']) and set(covered.values()) == {1}


def test_orbit_split_and_walk():
    sys.path.insert(0, CSRC)
    import gen_pip
    group = gen_pip.trimer_group()
    assert len(group) == 24 and len(set(group)) == 24
    gs = set(group)
    for h1 in group:  # closed under composition
        for h2 in group:
            assert gen_pip.compose(h1, h2) in gs
    a, b, mask = gen_pip.group_cycle(group)
    h, visited = tuple(range(36)), []
    for k in range(24):
        visited.append(h)
        h = gen_pip.compose(h, a if (mask >> k) & 1 else b)
    assert h == tuple(range(36)) and sorted(visited) == sorted(group)
    data = np.load(os.path.join(PARAMS, 'pip_monomials.npz'))
    with open(os.path.join(PARAMS, 'mbpol_tables.json')) as fh:
        coef = json.load(fh)['threebody']['thefit']
    qrows, qcoef, rrows, rcoef = gen_pip.orbit_loop_split(data['poly3b'], coef, group)
    # every monomial once: the passes cover each Q orbit |group| / |stabiliser| times over
    covered = {}
    for r in qrows:
        m = tuple(int(v) for v in r[1:] if v >= 0)
        orb = {tuple(sorted(hh[v] for v in m)) for hh in group}
        for o in orb:
            covered[o] = covered.get(o, 0) + 1
    for r in rrows:
        m = tuple(int(v) for v in r[1:] if v >= 0)
        covered[m] = covered.get(m, 0) + 1
    assert len(covered) == len(data['poly3b']) and set(covered.values()) == {1}


def water_pos_A(rng, n):
    sys.path.insert(0, ROOT)
    import mbpol_openmm_plugin_amd as m
    out = []
    for _ in range(n):
        w = m.systems.monomer(rng, rng.uniform(-3, 3, 3))
        # stretch and bend beyond the thermal range too
        w[1] = w[0] + (w[1] - w[0]) * rng.uniform(0.85, 1.2)
        w[2] = w[0] + (w[2] - w[0]) * rng.uniform(0.85, 1.2)
        out.append(w)
    return out


def test_water_terms_match_oracle(host, oracle):
    rng = np.random.default_rng(5)
    for w in water_pos_A(rng, 12):
        pos9 = np.ascontiguousarray(w.reshape(9))
        grad = np.zeros(9)
        e = host.host_water_onebody(_p(pos9), _p(grad))
        e_ref, f_ref = oracle.onebody(w * 0.1, np.array([[0, 1, 2]]))
        assert abs(e - e_ref) <= 1e-12 * max(1.0, abs(e_ref))
        f = -10.0 * grad.reshape(3, 3)  # kJ/mol/nm
        assert np.abs(f - f_ref).max() <= 1e-11 * np.abs(f_ref).max()
        q, dq = np.zeros(4), np.zeros(27)
        host.host_water_charges(_p(pos9), _p(q), _p(dq))
        pos4 = np.vstack([w, w[0]]) * 0.1  # the M position does not enter the charges
        q_ref, dq_ref = oracle.water_charges(pos4)
        np.testing.assert_allclose(q, q_ref, rtol=1e-12, atol=1e-15)
        dq_ref = dq_ref[:3].reshape(27)  # rows O, H1, H2; per nm
        assert np.abs(10.0 * dq - dq_ref).max() <= 1e-11 * np.abs(dq_ref).max()
