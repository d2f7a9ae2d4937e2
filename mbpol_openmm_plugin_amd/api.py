"""Python host layer of the MI355X MB-pol platform.

Mirrors the reference plugin's Force API (openmmapi/include/openmm/MBPol*Force.h)
and the slice of OpenMM's System / Context / State that drives it, on top of
the C ABI in include/mbpol_hip.h (libmbpol_hip.so, loaded with ctypes).
Method names, argument meaning, units (nm, kJ/mol, e, nm^3) and error
behaviour follow the reference; errors raised by the library surface as
``MBPolError`` (the reference throws OpenMMException).

There is no CPU fallback: creating a Context needs the HIP library and a GPU.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libmbpol_hip.so')

MBPOL_ONEBODY, MBPOL_TWOBODY, MBPOL_THREEBODY, MBPOL_DISPERSION, MBPOL_ELECTROSTATICS = 1, 2, 4, 8, 16
MBPOL_ALL = 31
COMPONENTS = ('onebody', 'twobody', 'threebody', 'dispersion', 'electrostatics')


class MBPolError(RuntimeError):
    """A failed library call (the reference raises OpenMMException)."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


class _EvalStats(ctypes.Structure):
    _fields_ = [('iterations', ctypes.c_int32), ('epsilon', ctypes.c_double), ('converged', ctypes.c_int32),
                ('num_dimers', ctypes.c_int64), ('num_trimers', ctypes.c_int64),
                ('num_elec_pairs', ctypes.c_int64), ('num_disp_pairs', ctypes.c_int64),
                ('num_pol_pairs', ctypes.c_int64)]


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def library():
    """Load libmbpol_hip.so (fails loudly when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MBPolError(-1, 'libmbpol_hip.so is not built (run __graft_entry__.build())')
        L = ctypes.CDLL(LIB_PATH)
        L.mbpol_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise MBPolError(rc, library().mbpol_last_error().decode())


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _pd(a):
    return a.ctypes.data_as(_dp)


def _pi(a):
    return a.ctypes.data_as(_ip)


def device_count():
    c = ctypes.c_int32(0)
    library().mbpol_device_count(ctypes.byref(c))
    return c.value


# ---------------------------------------------------------------------------
# Forces
# ---------------------------------------------------------------------------
class Force:
    def __init__(self):
        self._group = 0

    def setForceGroup(self, g):
        self._group = int(g)

    def getForceGroup(self):
        return self._group


class MBPolOneBodyForce(Force):
    """openmmapi/include/openmm/MBPolOneBodyForce.h:51-147."""
    NonPeriodic = 0
    Periodic = 1

    def __init__(self):
        super().__init__()
        self._mol = []
        self._method = self.NonPeriodic

    def addOneBody(self, particleIndices):
        self._mol.append([int(v) for v in particleIndices])
        return len(self._mol) - 1

    def getNumOneBodys(self):
        return len(self._mol)

    def getOneBodyParameters(self, index):
        return list(self._mol[index])

    def setOneBodyParameters(self, index, particleIndices):
        self._mol[index] = [int(v) for v in particleIndices]

    def getNonbondedMethod(self):
        return self._method

    def setNonbondedMethod(self, method):
        self._method = int(method)

    def updateParametersInContext(self, context):
        context._update_force(self)


class _NBodyForce(Force):
    NoCutoff = 0
    CutoffPeriodic = 1
    CutoffNonPeriodic = 2

    def __init__(self):
        super().__init__()
        self._mol = []
        self._method = self.NoCutoff
        self._cutoff = 1.0e10  # openmmapi/src/MBPolTwoBodyForce.cpp:43

    def addParticle(self, particleIndices):
        self._mol.append([int(v) for v in particleIndices])
        return len(self._mol) - 1

    def getNumMolecules(self):
        return len(self._mol)

    getNumParticles = getNumMolecules

    def getParticleParameters(self, index):
        return list(self._mol[index])

    def setParticleParameters(self, index, particleIndices):
        self._mol[index] = [int(v) for v in particleIndices]

    def getNonbondedMethod(self):
        return self._method

    def setNonbondedMethod(self, method):
        self._method = int(method)

    def getCutoff(self):
        return self._cutoff

    def setCutoff(self, cutoff):
        self._cutoff = float(cutoff)

    def updateParametersInContext(self, context):
        context._update_force(self)


class MBPolTwoBodyForce(_NBodyForce):
    """openmmapi/include/openmm/MBPolTwoBodyForce.h:57-178."""


class MBPolThreeBodyForce(_NBodyForce):
    """openmmapi/include/openmm/MBPolThreeBodyForce.h:57-178."""


class MBPolElectrostaticsForce(Force):
    """openmmapi/include/openmm/MBPolElectrostaticsForce.h:56-323; defaults from
    openmmapi/src/MBPolElectrostaticsForce.cpp:43-49."""
    NoCutoff = 0
    PME = 1

    def __init__(self):
        super().__init__()
        self._p = []
        self._method = self.NoCutoff
        self._cutoff = 0.9
        self._aewald = 0.0
        self._grid = [0, 0, 0]
        self._tol = 1e-4
        self._max_iter = 200
        self._eps = 1.0e-7
        self._redistribute = True
        self._thole = [0.4, 0.4, 0.055, 0.626, 0.055]

    def addElectrostatics(self, charge, moleculeIndex, atomType, dampingFactor, polarity):
        self._p.append([float(charge), int(moleculeIndex), int(atomType), float(dampingFactor), float(polarity)])
        return len(self._p) - 1

    def getNumElectrostatics(self):
        return len(self._p)

    def getElectrostaticsParameters(self, index):
        return tuple(self._p[index])

    def setElectrostaticsParameters(self, index, charge, moleculeIndex, atomType, dampingFactor, polarity):
        self._p[index] = [float(charge), int(moleculeIndex), int(atomType), float(dampingFactor), float(polarity)]

    def getNonbondedMethod(self):
        return self._method

    def setNonbondedMethod(self, method):
        self._method = int(method)

    def getCutoffDistance(self):
        return self._cutoff

    def setCutoffDistance(self, distance):
        self._cutoff = float(distance)

    def setIncludeChargeRedistribution(self, flag):
        self._redistribute = bool(flag)

    def getIncludeChargeRedistribution(self):
        return self._redistribute

    def setTholeParameters(self, thole):
        self._thole = [float(v) for v in thole]

    def getTholeParameters(self):
        return list(self._thole)

    def getAEwald(self):
        return self._aewald

    def setAEwald(self, aewald):
        self._aewald = float(aewald)

    def getPmeBSplineOrder(self):
        return 5

    def getPmeGridDimensions(self):
        return list(self._grid)

    def setPmeGridDimensions(self, dims):
        self._grid = [int(v) for v in dims]

    def getEwaldErrorTolerance(self):
        return self._tol

    def setEwaldErrorTolerance(self, tol):
        self._tol = float(tol)

    def getMutualInducedMaxIterations(self):
        return self._max_iter

    def setMutualInducedMaxIterations(self, n):
        self._max_iter = int(n)

    def getMutualInducedTargetEpsilon(self):
        return self._eps

    def setMutualInducedTargetEpsilon(self, eps):
        self._eps = float(eps)

    def updateParametersInContext(self, context):
        context._update_force(self)

    def _arrays(self):
        p = np.array(self._p, dtype=np.float64).reshape(-1, 5)
        return p[:, 0], p[:, 1].astype(np.int32), p[:, 2].astype(np.int32), p[:, 3], p[:, 4]


class MBPolDispersionForce(Force):
    """The damped-C6 dispersion that python/mbpol.xml:37-108 builds as an OpenMM
    CustomNonbondedForce, as a Force of its own: ``-C6 tt6 / r^6`` with
    Tang-Toennies damping, per-type C6/d6 tables, pairs inside one molecule and
    types with an all-zero C6 row excluded, truncated at the cutoff."""
    NoCutoff = 0
    CutoffPeriodic = 1
    CutoffNonPeriodic = 2

    def __init__(self):
        super().__init__()
        self._p = []
        self._method = self.NoCutoff
        self._cutoff = 1.0
        self._ntypes = 0
        self._c6 = None
        self._d6 = None

    def setTypeTables(self, numTypes, c6, d6):
        """c6[t1 + numTypes*t2] in kJ nm^6/mol, d6 in nm^-1 (Discrete2DFunction layout)."""
        self._ntypes = int(numTypes)
        self._c6 = [float(v) for v in c6]
        self._d6 = [float(v) for v in d6]

    def addParticle(self, typeIndex, moleculeIndex):
        self._p.append([int(typeIndex), int(moleculeIndex)])
        return len(self._p) - 1

    def getNumParticles(self):
        return len(self._p)

    def getNonbondedMethod(self):
        return self._method

    def setNonbondedMethod(self, method):
        self._method = int(method)

    def getCutoff(self):
        return self._cutoff

    def setCutoff(self, cutoff):
        self._cutoff = float(cutoff)


# ---------------------------------------------------------------------------
# System / Context / State
# ---------------------------------------------------------------------------
class ThreeParticleAverageSite:
    def __init__(self, p1, p2, p3, w1, w2, w3):
        self.particles = (int(p1), int(p2), int(p3))
        self.weights = (float(w1), float(w2), float(w3))


class System:
    def __init__(self):
        self._masses = []
        self._vsites = {}
        self._forces = []
        self._box = None
        self._groups = None

    def addParticle(self, mass):
        self._masses.append(float(mass))
        return len(self._masses) - 1

    def getNumParticles(self):
        return len(self._masses)

    def getParticleMass(self, i):
        return self._masses[i]

    def setVirtualSite(self, index, site):
        self._vsites[int(index)] = site

    def isVirtualSite(self, index):
        return int(index) in self._vsites

    def setDefaultPeriodicBoxVectors(self, a, b, c):
        a, b, c = (np.asarray(v, dtype=np.float64) for v in (a, b, c))
        if np.count_nonzero(a[1:]) or b[0] or b[2] or c[0] or c[1]:
            raise MBPolError(1, 'only rectangular boxes are supported (as in the reference)')
        self._box = [a[0], b[1], c[2]]

    def getDefaultPeriodicBoxVectors(self):
        b = self._box or [0.0, 0.0, 0.0]
        return [b[0], 0, 0], [0, b[1], 0], [0, 0, b[2]]

    def addForce(self, force):
        self._forces.append(force)
        return len(self._forces) - 1

    def getNumForces(self):
        return len(self._forces)

    def getForce(self, i):
        return self._forces[i]

    def removeForce(self, i):
        del self._forces[i]

    def setWaterGroups(self, water_start):
        """Independent groups (HIP extension): waters [water_start[g], water_start[g+1])
        form group g; nothing interacts across groups and energies are also reported per
        group (Context.getGroupEnergies).  None removes the groups."""
        self._groups = None if water_start is None else [int(v) for v in water_start]

    def computeVirtualSites(self, positions):
        pos = np.array(positions, dtype=np.float64).reshape(-1, 3)
        for s, v in self._vsites.items():
            pos[s] = sum(w * pos[p] for p, w in zip(v.particles, v.weights))
        return pos


class State:
    def __init__(self, energies, forces, stats):
        self._e = energies
        self._f = forces
        self.stats = stats

    def getPotentialEnergy(self):
        """kJ/mol."""
        return float(self._e[5])

    def getForces(self, asNumpy=True):
        """kJ/mol/nm, virtual-site forces folded onto their parents."""
        return self._f

    def getComponentEnergies(self):
        return {k: float(self._e[i]) for i, k in enumerate(COMPONENTS)}


def _build_native_system(system):
    L = library()
    n = system.getNumParticles()
    h = ctypes.c_void_p()
    _check(L.mbpol_system_create(ctypes.c_int32(n), ctypes.byref(h)))
    try:
        if system._box is not None:
            _check(L.mbpol_system_set_box(h, _pd(_d(system._box))))
        for s, v in sorted(system._vsites.items()):
            _check(L.mbpol_system_add_virtual_site(h, s, *v.particles, *(ctypes.c_double(w) for w in v.weights)))
        for f in system._forces:
            if isinstance(f, MBPolOneBodyForce):
                idx = _i(np.array(f._mol, dtype=np.int32).reshape(-1, 3))
                _check(L.mbpol_system_set_onebody(h, len(f._mol), _pi(idx), f._method))
            elif isinstance(f, (MBPolTwoBodyForce, MBPolThreeBodyForce)):
                idx = _i(np.array(f._mol, dtype=np.int32).reshape(-1, 3))
                fn = L.mbpol_system_set_twobody if isinstance(f, MBPolTwoBodyForce) else L.mbpol_system_set_threebody
                _check(fn(h, len(f._mol), _pi(idx), f._method, ctypes.c_double(f._cutoff)))
            elif isinstance(f, MBPolElectrostaticsForce):
                q, mol, typ, damp, pol = f._arrays()
                grid = _i(f._grid)
                thole = _d(f._thole)
                _check(L.mbpol_system_set_electrostatics(
                    h, len(f._p), _pd(_d(q)), _pi(_i(mol)), _pi(_i(typ)), _pd(_d(damp)), _pd(_d(pol)), f._method,
                    ctypes.c_double(f._cutoff), ctypes.c_double(f._aewald), _pi(grid), ctypes.c_double(f._tol),
                    f._max_iter, ctypes.c_double(f._eps), int(f._redistribute), _pd(thole)))
            elif isinstance(f, MBPolDispersionForce):
                p = np.array(f._p, dtype=np.int32).reshape(-1, 2)
                _check(L.mbpol_system_set_dispersion(h, len(f._p), _pi(_i(p[:, 0])), _pi(_i(p[:, 1])), f._ntypes,
                                                     _pd(_d(f._c6)), _pd(_d(f._d6)), f._method,
                                                     ctypes.c_double(f._cutoff)))
            else:
                raise MBPolError(1, 'unsupported force type %s' % type(f).__name__)
        if system._groups is not None:
            ws = _i(system._groups)
            _check(L.mbpol_system_set_groups(h, len(ws) - 1, _pi(ws)))
    except Exception:
        L.mbpol_system_destroy(h)
        raise
    return h


PRECISIONS = {'mixed': 0, 'double': 1, 'single-pip': 2}
SOLVERS = {'sor': 0, 'pcg': 1}


class Context:
    """A HIP context for one System on one GPU.

    precision: 'mixed' (FP64 polynomials and electrostatics, FP32 dispersion
    pair arithmetic, FP64 accumulation), 'double' (FP64 throughout) or
    'single-pip' (mixed with FP32 2-/3-body polynomials; their energies then
    carry relative errors of order 1e-3).

    The induced dipoles are solved by the reference's SOR iteration in
    'double' mode and by preconditioned conjugate gradients otherwise; both
    stop on the reference's convergence measure (setDipoleSolver changes it)."""

    def __init__(self, system, platform='HIP', precision='mixed', device=0):
        if platform != 'HIP':
            raise MBPolError(1, 'this package provides the HIP platform only')
        self.system = system
        if precision not in PRECISIONS:
            raise MBPolError(1, 'unknown precision %r (one of %s)' % (precision, ', '.join(PRECISIONS)))
        self._precision = PRECISIONS[precision]
        self._positions = None
        L = library()
        sysh = _build_native_system(system)
        self._h = ctypes.c_void_p()
        try:
            _check(L.mbpol_context_create(sysh, self._precision, int(device), ctypes.byref(self._h)))
        finally:
            L.mbpol_system_destroy(sysh)
        self._mask = MBPOL_ALL

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            library().mbpol_context_destroy(h)
            self._h = None

    def setPositions(self, positions):
        self._positions = self.system.computeVirtualSites(positions)

    def getPositions(self):
        return self._positions

    def computeVirtualSites(self):
        self._positions = self.system.computeVirtualSites(self._positions)

    def setPeriodicBoxVectors(self, a, b, c):
        box = _d([a[0], b[1], c[2]])
        _check(library().mbpol_context_set_box(self._h, _pd(box)))
        self.system._box = list(box)

    def setComponents(self, mask):
        self._mask = int(mask)

    def setDipoleGuess(self, mode, dipoles=None):
        modes = {'alpha_e': 0, 'zero': 1, 'previous': 2, 'user': 3, 'learned': 4}
        arr = _d(dipoles) if dipoles is not None else None
        _check(library().mbpol_context_set_dipole_guess(self._h, modes[mode], _pd(arr) if arr is not None else None))

    def setDipoleSolver(self, solver):
        """'sor' (the reference's iteration) or 'pcg' (conjugate gradients, same stopping rule)."""
        if solver not in SOLVERS:
            raise MBPolError(1, 'unknown dipole solver %r (one of %s)' % (solver, ', '.join(SOLVERS)))
        _check(library().mbpol_context_set_solver(self._h, SOLVERS[solver]))

    def getTuning(self, name):
        """Launch shape picked by the context's autotuning ('site_lanes'), a grid cap
        ('pip_blocks', 'pip3_blocks', 'disp_blocks'), the 2- and 3-body kernels ('pip2_kernel', 'pip3_kernel':
        0 the loop over the relabelling group, 1 term by term),
        a capacity ('row_capacity_pol', 'row_capacity_other', 'dimer_capacity',
        'lower_capacity', 'trimer_capacity'), the tuned overlap ('fork_early': 1 fork
        after the virtual sites, 0 after the electrostatics set-up, 2 after the virtual
        sites with the short-range kernels enqueued by a host thread; 'pip_waves',
        'pip3_waves', 'disp_waves'; 'tune_evaluations') or a graph-mode counter
        ('graph_steps', 'graph_captures', 'graph_replays', 'graph_fallbacks')."""
        v = ctypes.c_int32(0)
        _check(library().mbpol_context_get_tuning(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def setListMode(self, full_chain_list):
        _check(library().mbpol_context_set_list_mode(self._h, int(bool(full_chain_list))))

    def setGraph(self, enabled):
        """Graph mode (HIP extension): evaluations captured in a HIP graph and replayed."""
        _check(library().mbpol_context_set_graph(self._h, int(bool(enabled))))

    def getState(self, getEnergy=True, getForces=True, components=None):
        if self._positions is None:
            raise MBPolError(1, 'positions have not been set')
        pos = _d(self._positions)
        forces = np.zeros_like(pos)
        energies = np.zeros(6)
        mask = self._mask if components is None else int(components)
        _check(library().mbpol_context_evaluate(self._h, _pd(pos), _pd(forces), _pd(energies), ctypes.c_uint32(mask)))
        return State(energies, forces, self.getStats())

    # -- device-resident evaluation and kernel timing (HIP extensions, no reference counterpart) --
    def uploadPositions(self, positions):
        """Copy positions (virtual sites recomputed) into the context's device buffer."""
        self._positions = self.system.computeVirtualSites(positions)
        _check(library().mbpol_context_upload_positions(self._h, _pd(_d(self._positions))))

    def evaluateResident(self, components=None):
        """Evaluate on the uploaded positions; forces stay on the device. Returns the 6 energies (kJ/mol)."""
        e = np.zeros(6)
        mask = self._mask if components is None else int(components)
        _check(library().mbpol_context_evaluate_resident(self._h, _pd(e), ctypes.c_uint32(mask)))
        return e

    def downloadForces(self):
        """Forces (N x 3, kJ/mol/nm) of the last resident evaluation."""
        f = np.zeros((self.system.getNumParticles(), 3))
        _check(library().mbpol_context_download_forces(self._h, _pd(f)))
        return f

    def setTiming(self, enabled):
        """Enable (and reset) or disable HIP-event timing of phases and iteration kernels."""
        _check(library().mbpol_context_set_timing(self._h, int(bool(enabled))))

    def getTiming(self, name):
        """(accumulated milliseconds, number of timed spans) of a phase or kernel."""
        ms, cnt = ctypes.c_double(0), ctypes.c_int64(0)
        _check(library().mbpol_context_get_timing(self._h, name.encode(), ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    def getGroupEnergies(self):
        """(groups, 6) energies of the last evaluation, kJ/mol: 1B, 2B, 3B, dispersion,
        electrostatics, total of every group."""
        ng = len(self.system._groups) - 1
        out = np.zeros((ng, 6))
        _check(library().mbpol_context_get_group_energies(self._h, _pd(out)))
        return out

    def getStats(self):
        st = _EvalStats()
        library().mbpol_context_get_stats(self._h, ctypes.byref(st))
        return {k: getattr(st, k) for k, _ in _EvalStats._fields_}

    def getInducedDipoles(self):
        out = np.zeros((self.system.getNumParticles(), 3))
        _check(library().mbpol_context_get_dipoles(self._h, _pd(out)))
        return out

    def getCharges(self):
        out = np.zeros(self.system.getNumParticles())
        _check(library().mbpol_context_get_charges(self._h, _pd(out)))
        return out

    def getDimerList(self):
        c = ctypes.c_int64(0)
        L = library()
        _check(L.mbpol_context_get_dimer_list(self._h, None, ctypes.c_int64(0), ctypes.byref(c)))
        out = np.zeros((max(c.value, 1), 2), dtype=np.int32)
        _check(L.mbpol_context_get_dimer_list(self._h, _pi(out), ctypes.c_int64(c.value), ctypes.byref(c)))
        return out[:c.value]

    def getTrimerList(self):
        c = ctypes.c_int64(0)
        L = library()
        _check(L.mbpol_context_get_trimer_list(self._h, None, ctypes.c_int64(0), ctypes.byref(c)))
        out = np.zeros((max(c.value, 1), 3), dtype=np.int32)
        _check(L.mbpol_context_get_trimer_list(self._h, _pi(out), ctypes.c_int64(c.value), ctypes.byref(c)))
        return out[:c.value]

    def _update_force(self, force):
        if isinstance(force, MBPolElectrostaticsForce):
            q, mol, typ, damp, pol = force._arrays()
            _check(library().mbpol_context_update_electrostatics(self._h, len(q), _pd(_d(q)), _pi(_i(mol)),
                                                                 _pi(_i(typ)), _pd(_d(damp)), _pd(_d(pol))))
        elif isinstance(force, MBPolOneBodyForce):
            idx = _i([force.getOneBodyParameters(i) for i in range(force.getNumOneBodys())]).reshape(-1)
            _check(library().mbpol_context_update_onebody(self._h, force.getNumOneBodys(), _pi(idx)))
        elif isinstance(force, (MBPolTwoBodyForce, MBPolThreeBodyForce)):
            n = force.getNumMolecules()
            idx = _i([force.getParticleParameters(i) for i in range(n)]).reshape(-1)
            fn = (library().mbpol_context_update_twobody if isinstance(force, MBPolTwoBodyForce)
                  else library().mbpol_context_update_threebody)
            _check(fn(self._h, n, _pi(idx)))
        else:
            raise MBPolError(1, 'updateParametersInContext: unknown Force %s' % type(force).__name__)


def _context_get_potential(self, points):
    """MBPolElectrostaticsForce::getElectrostaticPotential: potential (kJ/mol/e) at
    points (nm) from the site charges and induced dipoles at the current positions.
    As the reference's kernel (MBPolReferenceKernels.cpp:260-283), it first runs the
    electrostatics set-up and dipole solve, here as an electrostatics-only evaluation."""
    pts = _d(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    out = np.zeros(len(pts))
    self.getState(getForces=False, components=MBPOL_ELECTROSTATICS)
    _check(library().mbpol_context_get_electrostatic_potential(self._h, len(pts), _pd(pts), _pd(out)))
    return out


def _context_get_moments(self):
    """MBPolElectrostaticsForce::getSystemElectrostaticsMoments: [net charge (e),
    dipole (3, Debye), traceless quadrupole (3 x 3 row-major, Buckingham)] about
    the System's centre of mass, at the current positions (the electrostatics are
    evaluated first, as the reference's kernel does, MBPolReferenceKernels.cpp:285-303)."""
    sysm = self.system
    masses = _d([sysm.getParticleMass(i) for i in range(sysm.getNumParticles())])
    out = np.zeros(13)
    self.getState(getForces=False, components=MBPOL_ELECTROSTATICS)
    _check(library().mbpol_context_get_system_electrostatics_moments(self._h, _pd(masses), _pd(out)))
    return out


Context.getElectrostaticPotential = _context_get_potential
Context.getSystemElectrostaticsMoments = _context_get_moments


def _context_set_dipole_convergence(self, target_epsilon, max_iterations=200):
    """setMutualInducedTargetEpsilon / setMutualInducedMaxIterations on the live context."""
    L = library()
    L.mbpol_context_set_dipole_convergence.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32]
    _check(L.mbpol_context_set_dipole_convergence(self._h, float(target_epsilon), int(max_iterations)))


Context.setDipoleConvergence = _context_set_dipole_convergence

_fp = ctypes.POINTER(ctypes.c_float)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _pf(a):
    return a.ctypes.data_as(_fp)


GUESS_PARAMS = 7491
GUESS_FEATURES = 48


class GuessModel:
    """Learned initial guess of the induced dipoles (HIP extension; the
    reference always starts from alpha * E_fixed).

    A per-atom network (48 radial features -> 64 -> 64 -> 3, SiLU; BF16 MFMA
    with FP32 accumulation) predicts (mu - alpha E_fixed) / |alpha E_fixed| in
    a frame of the atom anchored on its alpha E_fixed.
    Snapshots of converged dipoles are taken from the context's evaluations
    (addSnapshot after getState), the network is trained on the GPU (train),
    and attach() makes it the context's initial guess.  The solve still runs to
    the same tolerance, so the guess changes the iteration count only."""

    KINDS = {'alpha_e': 0, 'previous': 1}

    def __init__(self, context, max_snapshots, seed=1, kind='alpha_e'):
        """kind: 'alpha_e' -- the network corrects alpha E_fixed (anchored on it);
        'previous' -- it corrects the previous evaluation's converged dipoles,
        anchored on the change of alpha E_fixed since then (snapshots need two
        earlier evaluations of the context)."""
        if kind not in self.KINDS:
            raise MBPolError(1, 'unknown guess kind %r (one of %s)' % (kind, ', '.join(self.KINDS)))
        L = library()
        L.mbpol_guess_create_kind.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32,
                                              ctypes.c_void_p]
        self.context = context
        self.kind = kind
        self._h = ctypes.c_void_p()
        _check(L.mbpol_guess_create_kind(context._h, int(max_snapshots), int(seed), self.KINDS[kind],
                                         ctypes.byref(self._h)))
        self.natoms = 3 * (context.system.getNumParticles() // 4)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            L = library()
            L.mbpol_guess_destroy.argtypes = [ctypes.c_void_p]
            L.mbpol_guess_destroy(h)
            self._h = None

    def addSnapshot(self):
        _check(library().mbpol_guess_add_snapshot(self._h))

    def numSnapshots(self):
        c = ctypes.c_int32(0)
        _check(library().mbpol_guess_num_snapshots(self._h, ctypes.byref(c)))
        return c.value

    def getSnapshot(self, k):
        """(features [3 W, 48], targets [3 W, 3]) of stored snapshot k."""
        f = np.zeros((self.natoms, GUESS_FEATURES), dtype=np.float32)
        t = np.zeros((self.natoms, 3), dtype=np.float32)
        _check(library().mbpol_guess_get_snapshot(self._h, int(k), _pf(f), _pf(t)))
        return f, t

    def train(self, epochs, batch=4096, learning_rate=1e-3, ntrain=None):
        """Adam epochs over the first `ntrain` stored snapshots (all by default);
        returns the mean loss of each epoch."""
        L = library()
        L.mbpol_guess_train_range.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_double, _fp]
        out = np.zeros(max(int(epochs), 1), dtype=np.float32)
        n = self.numSnapshots() if ntrain is None else int(ntrain)
        _check(L.mbpol_guess_train_range(self._h, n, int(epochs), int(batch), float(learning_rate), _pf(out)))
        return out[:int(epochs)]

    def predictLast(self):
        """The network's guess (particles x 3) for the last evaluated configuration: from its
        alpha E_fixed, or (kind 'previous') from the evaluation before it."""
        out = np.zeros((self.context.system.getNumParticles(), 3))
        _check(library().mbpol_guess_predict_last(self._h, _pd(out)))
        return out

    def validationLoss(self, first, count, batch=4096):
        """Mean squared error (units of the output scale) on stored snapshots [first, first + count)."""
        L = library()
        L.mbpol_guess_validation_loss.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                  ctypes.POINTER(ctypes.c_double)]
        v = ctypes.c_double(0)
        _check(L.mbpol_guess_validation_loss(self._h, int(first), int(count), int(batch), ctypes.byref(v)))
        return v.value

    def attach(self, on=True):
        _check(library().mbpol_guess_attach(self._h, int(bool(on))))

    def getParameters(self):
        p = np.zeros(GUESS_PARAMS, dtype=np.float32)
        n = np.zeros(2 * GUESS_FEATURES + 1, dtype=np.float32)
        _check(library().mbpol_guess_get_params(self._h, _pf(p), _pf(n)))
        return p, n

    def setParameters(self, params, norm=None):
        p = _f32(params)
        n = _f32(norm) if norm is not None else None
        _check(library().mbpol_guess_set_params(self._h, _pf(p), _pf(n) if n is not None else None))

    def gradient(self, x, t):
        """(gradient, loss) of (1/n) sum |net(x_i) - t_i / scale|^2 as one training step computes it."""
        x, t = _f32(np.asarray(x).reshape(-1, GUESS_FEATURES)), _f32(np.asarray(t).reshape(-1, 3))
        g = np.zeros(GUESS_PARAMS, dtype=np.float32)
        loss = ctypes.c_float(0)
        _check(library().mbpol_guess_gradient(self._h, len(x), _pf(x), _pf(t), _pf(g), ctypes.byref(loss)))
        return g, loss.value

    def forward(self, x):
        x = _f32(np.asarray(x).reshape(-1, GUESS_FEATURES))
        y = np.zeros((len(x), 3), dtype=np.float32)
        _check(library().mbpol_guess_forward(self._h, len(x), _pf(x), _pf(y)))
        return y

    def features(self):
        """(features [3 W, 48], frames [3 W, 3, 3]) of the context's last evaluation."""
        f = np.zeros((self.natoms, GUESS_FEATURES), dtype=np.float32)
        r = np.zeros((self.natoms, 3, 3), dtype=np.float32)
        _check(library().mbpol_guess_features(self._h, _pf(f), _pf(r)))
        return f, r


class LangevinIntegrator:
    """Langevin dynamics (BAOAB splitting) on the context's device-resident
    positions, with forces from the same context (HIP extension; mirrors
    OpenMM's LangevinIntegrator(temperature K, friction 1/ps, step ps)).
    Virtual sites (mass 0) follow their parents.  `components` restricts the
    forces (and energies) to a subset of the Forces."""

    def __init__(self, context, temperature, friction, step_size, seed=1, components=MBPOL_ALL):
        L = library()
        L.mbpol_md_create.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p]
        self.context = context
        sysm = context.system
        masses = _d([0.0 if sysm.isVirtualSite(i) else sysm.getParticleMass(i)
                     for i in range(sysm.getNumParticles())])
        if context.getPositions() is None:
            raise MBPolError(1, 'positions have not been set')
        context.uploadPositions(context.getPositions())
        self._h = ctypes.c_void_p()
        _check(L.mbpol_md_create(context._h, _pd(masses), float(temperature), float(friction), float(step_size),
                                 int(seed), int(components), ctypes.byref(self._h)))
        self._n = len(masses)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            L = library()
            L.mbpol_md_destroy.argtypes = [ctypes.c_void_p]
            L.mbpol_md_destroy(h)
            self._h = None

    def step(self, steps):
        """Run steps; returns (energies[6] of the last force evaluation, kinetic energy), kJ/mol."""
        e = np.zeros(6)
        ke = ctypes.c_double(0)
        _check(library().mbpol_md_step(self._h, int(steps), _pd(e), ctypes.byref(ke)))
        return e, ke.value

    def getPositions(self):
        out = np.zeros((self._n, 3))
        _check(library().mbpol_md_get_positions(self._h, _pd(out)))
        return out
