"""ctypes binding of the FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
never by the product package.  Units: nm, kJ/mol, kJ/mol/nm.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'build', 'liboracle.so')
REF_PATH = os.path.join(HERE, '_ref', 'libref_poly.so')

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class ElecParams(ctypes.Structure):
    _fields_ = [('method', ctypes.c_int), ('cutoff', ctypes.c_double), ('alpha', ctypes.c_double),
                ('grid', ctypes.c_int * 3), ('thole', ctypes.c_double * 5), ('max_iter', ctypes.c_int),
                ('target_eps', ctypes.c_double), ('charge_redistribution', ctypes.c_int),
                ('box', ctypes.c_double * 3)]


class ElecStats(ctypes.Structure):
    _fields_ = [('iterations', ctypes.c_int), ('epsilon', ctypes.c_double), ('converged', ctypes.c_int),
                ('setup_s', ctypes.c_double), ('iteration_s', ctypes.c_double), ('energy_s', ctypes.c_double)]


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _pd(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _pi(a):
    return a.ctypes.data_as(_ip) if a is not None else None


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError('oracle library not built: run `make -C oracle`')
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_pip2b.restype = ctypes.c_double
        L.oracle_pip3b.restype = ctypes.c_double
        L.oracle_onebody.restype = ctypes.c_double
        L.oracle_dimer_list.restype = ctypes.c_int64
        L.oracle_dimer_list.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double, _ip, ctypes.c_int64]
        L.oracle_trimer_list.restype = ctypes.c_int64
        L.oracle_trimer_list.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _ip,
                                         ctypes.c_int64]
        L.oracle_twobody.restype = ctypes.c_double
        L.oracle_threebody.restype = ctypes.c_double
        L.oracle_threebody.argtypes = [ctypes.c_int64, _ip, _dp, _ip, ctypes.c_int, _dp, _dp]
        L.oracle_thole_scale.restype = ctypes.c_double
        L.oracle_thole_scale.argtypes = [ctypes.POINTER(ElecParams), ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int]
        L.oracle_gammq.restype = ctypes.c_double
        L.oracle_gammq.argtypes = [ctypes.c_double, ctypes.c_double]
        L.oracle_electrostatic_potential.restype = ctypes.c_int
        L.oracle_system_moments.restype = ctypes.c_int
        L.oracle_dispersion.restype = ctypes.c_double
        L.oracle_dispersion.argtypes = [ctypes.c_int, _dp, _ip, _ip, ctypes.c_int, _dp, _dp, ctypes.c_int,
                                        ctypes.c_double, _dp, _dp]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own polynomial/gammq sources compiled into oracle/_ref (may be absent)."""
    if not os.path.exists(REF_PATH):
        return None
    L = ctypes.CDLL(REF_PATH)
    for f in ('ref_poly_2b', 'ref_poly_3b', 'ref_gammq', 'ref_gammln'):
        getattr(L, f).restype = ctypes.c_double
    L.ref_gammq.argtypes = [ctypes.c_double, ctypes.c_double]
    L.ref_gammln.argtypes = [ctypes.c_double]
    return L


def pip2b(x):
    x = _d(x)
    g = np.zeros(31)
    e = lib().oracle_pip2b(_pd(x), _pd(g))
    return e, g


def pip3b(x):
    x = _d(x)
    g = np.zeros(36)
    e = lib().oracle_pip3b(_pd(x), _pd(g))
    return e, g


def _box(box):
    return _d(box if box is not None else np.zeros(3))


def onebody(pos, idx, periodic=False, box=None):
    pos, idx = _d(pos), _i(idx)
    f = np.zeros_like(pos)
    e = lib().oracle_onebody(len(idx), _pd(pos), _pi(idx), int(periodic), _pd(_box(box)), _pd(f))
    return e, f


def dimer_list(opos, rc, periodic=False, box=None):
    opos = _d(opos)
    n = len(opos)
    cap = max(1, n * (n - 1) // 2)
    out = np.zeros((cap, 2), np.int32)
    c = lib().oracle_dimer_list(n, _pd(opos), _pd(_box(box)), int(periodic), rc, _pi(out), cap)
    return out[:c].copy()


def trimer_list(opos, rc, periodic=False, box=None, rmin=0.0):
    opos = _d(opos)
    n = len(opos)
    c = lib().oracle_trimer_list(n, _pd(opos), _pd(_box(box)), int(periodic), rc, rmin, None, 0)
    out = np.zeros((max(c, 1), 3), np.int32)
    lib().oracle_trimer_list(n, _pd(opos), _pd(_box(box)), int(periodic), rc, rmin, _pi(out), c)
    return out[:c].copy()


def twobody(pos, idx, pairs, periodic=False, box=None):
    pos, idx, pairs = _d(pos), _i(idx), _i(pairs)
    f = np.zeros_like(pos)
    e = lib().oracle_twobody(len(pairs), _pi(pairs), _pd(pos), _pi(idx), int(periodic), _pd(_box(box)), _pd(f))
    return e, f


def threebody(pos, idx, trip, periodic=False, box=None):
    pos, idx, trip = _d(pos), _i(idx), _i(trip)
    f = np.zeros_like(pos)
    e = lib().oracle_threebody(len(trip), _pi(trip), _pd(pos), _pi(idx), int(periodic), _pd(_box(box)), _pd(f))
    return e, f


def elec_params(method=0, cutoff=0.9, alpha=0.0, grid=(0, 0, 0), thole=(0.4, 0.4, 0.055, 0.626, 0.055),
                max_iter=200, target_eps=1e-7, charge_redistribution=True, box=(0, 0, 0)):
    p = ElecParams()
    p.method, p.cutoff, p.alpha = method, cutoff, alpha
    p.grid[:] = list(grid)
    p.thole[:] = list(thole)
    p.max_iter, p.target_eps, p.charge_redistribution = max_iter, target_eps, int(charge_redistribution)
    p.box[:] = list(box)
    return p


def electrostatics(params, pos, charge, mol, atype, damp, pol, want_field=False):
    pos = _d(pos)
    n = len(pos)
    f = np.zeros_like(pos)
    dip = np.zeros_like(pos)
    fld = np.zeros_like(pos) if want_field else None
    e = ctypes.c_double(0)
    st = ElecStats()
    err = ctypes.create_string_buffer(256)
    rc = lib().oracle_electrostatics(ctypes.byref(params), n, _pd(pos), _pd(_d(charge)), _pi(_i(mol)), _pi(_i(atype)),
                                     _pd(_d(damp)), _pd(_d(pol)), _pd(f), ctypes.byref(e), ctypes.byref(st), _pd(dip),
                                     _pd(fld), err, 256)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    out = dict(energy=e.value, forces=f, dipoles=dip, iterations=st.iterations, epsilon=st.epsilon,
               setup_s=st.setup_s, iteration_s=st.iteration_s, energy_s=st.energy_s)
    if want_field:
        out['fixed_field'] = fld
    return out


def electrostatic_potential(params, pos, charge, mol, atype, damp, pol, points):
    """getElectrostaticPotential (MBPolReferenceElectrostaticsForce.cpp:1023-1086), kJ/mol/e."""
    pos, pts = _d(pos), _d(np.asarray(points).reshape(-1, 3))
    out = np.zeros(len(pts))
    rc = lib().oracle_electrostatic_potential(ctypes.byref(params), len(pos), _pd(pos), _pd(_d(charge)),
                                              _pi(_i(mol)), _pi(_i(atype)), _pd(_d(damp)), _pd(_d(pol)),
                                              len(pts), _pd(pts), _pd(out))
    if rc != 0:
        raise RuntimeError('induced dipoles did not converge')
    return out


def system_moments(params, pos, charge, mol, atype, damp, pol, masses):
    """calculateMBPolSystemElectrostaticsMoments (MBPolReferenceElectrostaticsForce.cpp:923-1021)."""
    pos = _d(pos)
    out = np.zeros(13)
    rc = lib().oracle_system_moments(ctypes.byref(params), len(pos), _pd(pos), _pd(_d(charge)), _pi(_i(mol)),
                                     _pi(_i(atype)), _pd(_d(damp)), _pd(_d(pol)), _pd(_d(masses)), _pd(out))
    if rc != 0:
        raise RuntimeError('induced dipoles did not converge')
    return out


def thole_scale(params, damp_i, damp_j, r, order, itype, mol=(0, 1), types=(0, 0), just_scale=False):
    return lib().oracle_thole_scale(ctypes.byref(params), damp_i, damp_j, mol[0], mol[1], types[0], types[1], r,
                                    int(just_scale), order, itype)


def water_charges(pos4):
    pos4 = _d(pos4)
    q = np.zeros(4)
    dq = np.zeros((4, 3, 3))
    lib().oracle_water_charges(_pd(pos4), _pd(q), _pd(dq))
    return q, dq


def solve_given_field(params, pos, mol, atype, damp, pol, field):
    pos = _d(pos)
    dip = np.zeros_like(pos)
    st = ElecStats()
    rc = lib().oracle_solve_given_field(ctypes.byref(params), len(pos), _pd(pos), _pi(_i(mol)), _pi(_i(atype)),
                                        _pd(_d(damp)), _pd(_d(pol)), _pd(_d(field)), _pd(dip), ctypes.byref(st))
    return dip, st.iterations, st.epsilon, rc == 0


def gammq(a, x):
    return lib().oracle_gammq(a, x)


def dispersion(pos, types, mol, c6, d6, method=0, cutoff=1e3, box=None):
    pos = _d(pos)
    c6, d6 = _d(c6), _d(d6)
    ntypes = int(round(len(c6) ** 0.5))
    f = np.zeros_like(pos)
    e = lib().oracle_dispersion(len(pos), _pd(pos), _pi(_i(types)), _pi(_i(mol)), ntypes, _pd(c6), _pd(d6), method,
                                cutoff, _pd(_box(box)), _pd(f))
    return e, f


def fft3d(grid, sign):
    g = np.ascontiguousarray(grid, dtype=np.complex128).copy()
    nx, ny, nz = g.shape
    lib().oracle_fft3d(g.ctypes.data_as(_dp), nx, ny, nz, sign)
    return g


# ---------------------------------------------------------------------------
# whole-System evaluation: the sum over the Forces of a System as the reference
# Context would compute it (virtual sites placed, their forces folded back).
# ---------------------------------------------------------------------------
COMPONENTS = ('onebody', 'twobody', 'threebody', 'dispersion', 'electrostatics')


def pme_parameters(cutoff, tol, box):
    """NonbondedForceImpl::calcPMEParameters (OpenMM 6.3-7.x; called at
    MBPolReferenceKernels.cpp:186-199): alpha = sqrt(-ln(2 tol))/rc and
    n = max(ceil(2 alpha L / (3 tol^(1/5))), 6)."""
    import math
    alpha = (1.0 / cutoff) * math.sqrt(-math.log(2.0 * tol))
    grid = [max(int(math.ceil(2 * alpha * L / (3 * tol ** 0.2))), 6) for L in box]
    return alpha, grid


def _elec_force_params(system):
    """The System's electrostatics Force and the oracle parameters it describes."""
    f = [f for f in system._forces if type(f).__name__ == 'MBPolElectrostaticsForce'][0]
    box = np.array(system._box) if system._box is not None else None
    method = f.getNonbondedMethod()
    alpha, grid = f.getAEwald(), f.getPmeGridDimensions()
    if method == 1 and (grid[0] == 0 or alpha == 0.0):
        alpha, grid = pme_parameters(f.getCutoffDistance(), f.getEwaldErrorTolerance(), box)
    P = elec_params(method=method, cutoff=f.getCutoffDistance(), alpha=alpha, grid=grid,
                    thole=f.getTholeParameters(), max_iter=f.getMutualInducedMaxIterations(),
                    target_eps=f.getMutualInducedTargetEpsilon(),
                    charge_redistribution=f.getIncludeChargeRedistribution(),
                    box=box if box is not None else (0, 0, 0))
    return f, P


def system_electrostatic_potential(system, positions, points):
    f, P = _elec_force_params(system)
    q, mol, typ, damp, pol = f._arrays()
    return electrostatic_potential(P, system.computeVirtualSites(positions), q, mol, typ, damp, pol, points)


def system_electrostatics_moments(system, positions):
    f, P = _elec_force_params(system)
    q, mol, typ, damp, pol = f._arrays()
    masses = [system.getParticleMass(i) for i in range(system.getNumParticles())]
    return system_moments(P, system.computeVirtualSites(positions), q, mol, typ, damp, pol, masses)


def evaluate_system(system, positions, components=31, return_lists=False):
    """Energies (dict, kJ/mol) and forces (N x 3, kJ/mol/nm) of a System built with
    mbpol_openmm_plugin_amd.api (read through its Force getters only)."""
    pos = system.computeVirtualSites(positions)
    n = len(pos)
    F = np.zeros((n, 3))
    E = dict.fromkeys(COMPONENTS, 0.0)
    box = np.array(system._box) if system._box is not None else None
    lists = {}
    for f in system._forces:
        name = type(f).__name__
        if name == 'MBPolOneBodyForce' and components & 1:
            idx = np.array([f.getOneBodyParameters(i) for i in range(f.getNumOneBodys())])
            e, ff = onebody(pos, idx, periodic=f.getNonbondedMethod() == 1, box=box)
            E['onebody'] += e
            F += ff
        elif name in ('MBPolTwoBodyForce', 'MBPolThreeBodyForce'):
            two = name == 'MBPolTwoBodyForce'
            if not components & (2 if two else 4):
                continue
            idx = np.array([f.getParticleParameters(i) for i in range(f.getNumMolecules())])
            periodic = f.getNonbondedMethod() == 1
            opos = pos[idx[:, 0]]
            rc = f.getCutoff()
            if two:
                pairs = dimer_list(opos, rc, periodic, box)
                e, ff = twobody(pos, idx, pairs, periodic, box)
                E['twobody'] += e
                lists['dimers'] = pairs
            else:
                trip = trimer_list(opos, rc, periodic, box)
                e, ff = threebody(pos, idx, trip, periodic, box)
                E['threebody'] += e
                lists['trimers'] = trip
            F += ff
        elif name == 'MBPolElectrostaticsForce' and components & 16:
            q, mol, typ, damp, pol = f._arrays()
            method = f.getNonbondedMethod()
            alpha, grid = f.getAEwald(), f.getPmeGridDimensions()
            if method == 1 and (grid[0] == 0 or alpha == 0.0):
                alpha, grid = pme_parameters(f.getCutoffDistance(), f.getEwaldErrorTolerance(), box)
            P = elec_params(method=method, cutoff=f.getCutoffDistance(), alpha=alpha, grid=grid,
                            thole=f.getTholeParameters(), max_iter=f.getMutualInducedMaxIterations(),
                            target_eps=f.getMutualInducedTargetEpsilon(),
                            charge_redistribution=f.getIncludeChargeRedistribution(),
                            box=box if box is not None else (0, 0, 0))
            r = electrostatics(P, pos, q, mol, typ, damp, pol)
            E['electrostatics'] += r['energy']
            F += r['forces']
            lists['elec'] = r
        elif name == 'MBPolDispersionForce' and components & 8:
            p = np.array(f._p).reshape(-1, 2)
            e, ff = dispersion(pos, p[:, 0], p[:, 1], f._c6, f._d6, method=f.getNonbondedMethod(),
                               cutoff=f.getCutoff(), box=box)
            E['dispersion'] += e
            F += ff
    for s, v in system._vsites.items():
        for p, w in zip(v.particles, v.weights):
            F[p] += w * F[s]
        F[s] = 0.0
    E['total'] = sum(E[k] for k in COMPONENTS)
    if return_lists:
        return E, F, lists
    return E, F
