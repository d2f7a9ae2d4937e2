"""The C-ABI boundary beyond the energy: the reference kernel interface's other
entry points and its error paths, through the library (HIP path) against the
FP64 oracle.

* getElectrostaticPotential (MBPolReferenceElectrostaticsForce.cpp:1023-1086)
  and getSystemElectrostaticsMoments (:923-1021; glue MBPolReferenceKernels.cpp:
  260-303) on the 14-water cluster (direct sum) and the 256-water PME box; the
  oracle functions are pinned by the multipole expansion in
  tests/test_oracle_pins.py.  FP64 path: 1e-9 relative; mixed: 1e-6.
* copyParametersToContext of the 1-, 2- and 3-body Forces
  (MBPolReferenceKernels.cpp:116-127,401-414,485-498): a changed index list is
  used by the next evaluation (checked against the oracle reading the same
  Forces), and a changed count is refused with MBPOL_ERR_COUNT_CHANGED.
* Every reference error forced from the host: non-convergence (max 1 dipole
  iteration) -> NOT_CONVERGED, after which a normal evaluation on the same
  context equals a fresh context's; a box below twice a cutoff ->
  BOX_TOO_SMALL; a cutoff above half the box at creation -> CUTOFF_TOO_LARGE;
  an electrostatics update with another count -> COUNT_CHANGED.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NOT_CONVERGED, BOX_TOO_SMALL, CUTOFF_TOO_LARGE, COUNT_CHANGED = 3, 4, 5, 6


@pytest.fixture(scope='module')
def mb():
    import mbpol_openmm_plugin_amd as m
    if m.device_count() < 1:
        pytest.fail('no GPU visible to libmbpol_hip.so')
    return m


def cluster14(mb):
    sysm = mb.forcefield.create_system(14, nonbonded_method='CutoffNonPeriodic', twobody_cutoff=1e10,
                                       threebody_cutoff=1e10, nonbonded_cutoff=1e3, dipole_epsilon=1e-8)
    return sysm, mb.forcefield.water_positions_with_m(mb.systems.config0())


def box256(mb):
    waters, L = mb.systems.config1()
    sysm = mb.forcefield.create_system(256, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5,
                                       dipole_epsilon=1e-8)
    return sysm, mb.forcefield.water_positions_with_m(waters), L


@pytest.mark.parametrize('precision', ['double', 'mixed'])
@pytest.mark.parametrize('which', ['cluster14', 'box256'])
def test_potential_and_moments(mb, oracle, precision, which):
    if which == 'cluster14':
        sysm, pos = cluster14(mb)
        lo, hi = pos.min(axis=0) - 0.3, pos.max(axis=0) + 0.3
    else:
        sysm, pos, L = box256(mb)
        lo, hi = np.zeros(3), np.full(3, L)
    rng = np.random.default_rng(17)
    pts = rng.uniform(lo, hi, (24, 3))
    # keep the points off the sites (the potential is singular there)
    d = np.linalg.norm(pts[:, None, :] - pos[None, :, :], axis=-1).min(axis=1)
    pts = pts[d > 0.08]
    ctx = mb.Context(sysm, precision=precision)
    ctx.setPositions(pos)
    phi = ctx.getElectrostaticPotential(pts)
    mom = ctx.getSystemElectrostaticsMoments()
    phi_ref = oracle.system_electrostatic_potential(sysm, pos, pts)
    mom_ref = oracle.system_electrostatics_moments(sysm, pos)
    tol = 1e-9 if precision == 'double' else 1e-6
    perr = np.abs(phi - phi_ref).max() / np.abs(phi_ref).max()
    derr = np.abs(mom[1:4] - mom_ref[1:4]).max() / np.abs(mom_ref[1:4]).max()
    qerr = np.abs(mom[4:] - mom_ref[4:]).max() / np.abs(mom_ref[4:]).max()
    print(which, precision, len(pts), 'points: potential err %.2e, dipole err %.2e, quadrupole err %.2e' %
          (perr, derr, qerr), 'net charge', mom[0], flush=True)
    assert perr <= tol and derr <= tol and qerr <= tol
    assert abs(mom[0] - mom_ref[0]) <= 1e-9


def test_update_molecule_forces(mb, oracle):
    waters = mb.systems.cluster(3, seed=5)
    sysm = mb.forcefield.create_system(3, nonbonded_method='CutoffNonPeriodic', twobody_cutoff=1e10,
                                       threebody_cutoff=1e10, nonbonded_cutoff=1e3, dipole_epsilon=1e-8,
                                       components=1 | 2 | 4)
    pos = mb.forcefield.water_positions_with_m(waters)
    ctx = mb.Context(sysm, precision='double')
    ctx.setPositions(pos)
    e0 = ctx.getState(components=1 | 2 | 4).getComponentEnergies()
    forces = {type(f).__name__: f for f in sysm._forces}
    # the 2-body Force's molecule 2 becomes a second copy of water 1: dimer (1, 2) then sits at
    # r_OO = 0 (no energy) and (0, 2) repeats (0, 1)
    f2 = forces['MBPolTwoBodyForce']
    f2.setParticleParameters(2, f2.getParticleParameters(1))
    f2.updateParametersInContext(ctx)
    # the 1-body Force: water 0 listed with its hydrogens swapped (the surface is symmetric)
    f1 = forces['MBPolOneBodyForce']
    o, h1, h2 = f1.getOneBodyParameters(0)
    f1.setOneBodyParameters(0, [o, h2, h1])
    f1.updateParametersInContext(ctx)
    st = ctx.getState(components=1 | 2 | 4)
    E, F = oracle.evaluate_system(sysm, pos, components=1 | 2 | 4)
    got = st.getComponentEnergies()
    print('before', e0, 'after', got, 'oracle', E, flush=True)
    assert abs(got['twobody'] - e0['twobody']) > 1e-3 * abs(e0['twobody'])
    for k in ('onebody', 'twobody', 'threebody'):
        assert abs(got[k] - E[k]) <= 1e-9 * max(abs(E[k]), 1.0), k
    assert np.abs(st.getForces() - F).max() <= 1e-8 * np.abs(F).max()
    # the count must not change
    two = mb.api._i(np.array([[0, 1, 2], [4, 5, 6]], dtype=np.int32).reshape(-1))
    assert mb.api.library().mbpol_context_update_threebody(ctx._h, 2, mb.api._pi(two)) == COUNT_CHANGED


@pytest.mark.parametrize('precision', ['double', 'mixed'])
def test_not_converged_then_recovers(mb, precision):
    sysm, pos, L = box256(mb)
    ctx = mb.Context(sysm, precision=precision)
    ctx.setPositions(pos)
    ctx.setDipoleConvergence(1e-8, 1)
    with pytest.raises(mb.MBPolError) as ei:
        ctx.getState()
    assert ei.value.code == NOT_CONVERGED
    ctx.setDipoleConvergence(1e-8, 200)
    a = ctx.getState()
    fresh = mb.Context(sysm, precision=precision)
    fresh.setPositions(pos)
    b = fresh.getState()
    print(precision, a.getPotentialEnergy(), b.getPotentialEnergy(), a.stats['iterations'], b.stats['iterations'],
          flush=True)
    assert abs(a.getPotentialEnergy() - b.getPotentialEnergy()) <= 1e-12 * abs(b.getPotentialEnergy())
    assert np.abs(a.getForces() - b.getForces()).max() <= 1e-10 * np.abs(b.getForces()).max()


def test_box_and_cutoff_errors(mb):
    waters, L = mb.systems.config1()
    pos = mb.forcefield.water_positions_with_m(waters)
    sysm = mb.forcefield.create_system(256, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5)
    ctx = mb.Context(sysm, precision='double')
    ctx.setPositions(pos)
    ctx.getState()
    # the electrostatics cutoff is 0.9 nm: a 1.7 nm box is below twice it
    with pytest.raises(mb.MBPolError) as ei:
        ctx.setPeriodicBoxVectors([1.7, 0, 0], [0, 1.7, 0], [0, 0, 1.7])
    assert ei.value.code == BOX_TOO_SMALL
    # at creation, a cutoff above half the box
    small = mb.forcefield.create_system(8, nonbonded_method='PME', box=1.7, ewald_error_tolerance=1e-5)
    with pytest.raises(mb.MBPolError) as ei:
        mb.Context(small, precision='double')
    assert ei.value.code == CUTOFF_TOO_LARGE


def test_count_changed_electrostatics(mb):
    sysm, pos = cluster14(mb)
    ctx = mb.Context(sysm, precision='double')
    ctx.setPositions(pos)
    ctx.getState()
    n = sysm.getNumParticles() - 4
    L = mb.api.library()
    z = mb.api._d(np.zeros(n))
    zi = mb.api._i(np.zeros(n, dtype=np.int32))
    rc = L.mbpol_context_update_electrostatics(ctx._h, n, mb.api._pd(z), mb.api._pi(zi), mb.api._pi(zi),
                                               mb.api._pd(z), mb.api._pd(z))
    assert rc == COUNT_CHANGED
