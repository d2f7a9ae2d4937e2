"""HIP path (libmbpol_hip.so through its C ABI) against the FP64 oracle.

Tolerances: the FP64 GPU path must agree with the oracle to 1e-9 relative
energy and 1e-8 relative RMS force (only summation order differs); the
mixed-precision path (FP32 dispersion pair arithmetic, FP64 accumulation)
to <= 1e-6 of the total energy per component and <= 1e-4 relative RMS force.
The FP32-polynomial mode is checked to its documented error (<= 1e-2 of each
2-/3-body energy) and its error is printed.  Neighbour lists must be
identical as sets."""
import os

import numpy as np
import pytest

import kat_data as K

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


@pytest.fixture(scope='module')
def mb():
    import mbpol_openmm_plugin_amd as m
    if m.device_count() < 1:
        pytest.fail('no GPU visible to libmbpol_hip.so')
    return m


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-30))


TOL = {'double': (1e-9, 1e-8), 'mixed': (1e-6, 1e-4)}


def compare(mb, oracle, sys, pos, precision, components=31, label='', solver=None, tol=None):
    ctx = mb.Context(sys, precision=precision)
    if solver is not None:
        ctx.setDipoleSolver(solver)
    ctx.setPositions(pos)
    st = ctx.getState(components=components)
    E, F = oracle.evaluate_system(sys, pos, components=components)
    got = st.getComponentEnergies()
    etol, ftol = tol or TOL[precision]
    scale = max(abs(E['total']), 1.0)
    report = {k: (got[k], E[k]) for k in got}
    print(label, precision, {k: '%.9g vs %.9g' % v for k, v in report.items()}, 'stats', st.stats, flush=True)
    for k in got:
        assert abs(got[k] - E[k]) <= etol * scale, (label, k, got[k], E[k])
    assert abs(st.getPotentialEnergy() - E['total']) <= etol * scale
    err = rel_rms(st.getForces(), F)
    print(label, precision, 'force rel rms err %.3e' % err, flush=True)
    assert err <= ftol, (label, err)
    return ctx, st


@pytest.mark.parametrize('precision', ['double', 'mixed'])
def test_kat_waters_onebody_twobody_threebody(mb, oracle, precision):
    waters = K.WATERS_NM
    sys = mb.forcefield.create_system(3, nonbonded_method='CutoffNonPeriodic', twobody_cutoff=10.0,
                                      threebody_cutoff=10.0, components=1 | 2 | 4)
    pos = mb.forcefield.water_positions_with_m(waters)
    ctx, st = compare(mb, oracle, sys, pos, precision, components=7, label='kat3')
    e = st.getComponentEnergies()
    assert abs(e['onebody'] / K.KCAL - 0.55975882 - 0) < 5  # three monomers: finite and close to the KAT scale


@pytest.mark.parametrize('precision', ['double', 'mixed'])
def test_config0_cluster_all_components(mb, oracle, precision):
    waters = mb.systems.config0()
    sys = mb.forcefield.create_system(14, nonbonded_method='CutoffNonPeriodic', twobody_cutoff=1e10,
                                      threebody_cutoff=1e10, nonbonded_cutoff=1e3, dipole_epsilon=1e-8)
    pos = mb.forcefield.water_positions_with_m(waters)
    compare(mb, oracle, sys, pos, precision, label='config0')


@pytest.mark.parametrize('precision', ['double', 'mixed'])
def test_config1_pbc_all_components(mb, oracle, precision):
    waters, L = mb.systems.config1()
    sys = mb.forcefield.create_system(256, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5,
                                      dipole_epsilon=1e-8)
    pos = mb.forcefield.water_positions_with_m(waters)
    compare(mb, oracle, sys, pos, precision, label='config1')


@pytest.mark.parametrize('case', ['config0', 'config1'])
def test_dipole_solvers(mb, oracle, case):
    """Conjugate gradients against the reference's SOR iteration, both in FP64
    and both stopped by the reference's measure at 1e-8.  The two stop at
    different points within that distance of the exact solution, and the
    energy expression (induction energy as -1/2 mu.E_fixed, exact only at
    convergence) is first order in the dipoles, so the energies differ at the
    1e-8 relative level (1.6e-8 measured on the 14-water cluster): checked at
    1e-7 of the energy and 1e-6 RMS force.  PCG needs at most half the field
    evaluations."""
    if case == 'config0':
        waters = mb.systems.config0()
        sys = mb.forcefield.create_system(14, nonbonded_method='CutoffNonPeriodic', twobody_cutoff=1e10,
                                          threebody_cutoff=1e10, nonbonded_cutoff=1e3, dipole_epsilon=1e-8,
                                          components=16)
    else:
        waters, L = mb.systems.config1()
        sys = mb.forcefield.create_system(256, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5,
                                          dipole_epsilon=1e-8, components=16)
    pos = mb.forcefield.water_positions_with_m(waters)
    _, st_pcg = compare(mb, oracle, sys, pos, 'double', components=16, label=case + ' pcg', solver='pcg',
                        tol=(1e-7, 1e-6))
    _, st_sor = compare(mb, oracle, sys, pos, 'double', components=16, label=case + ' sor', solver='sor')
    it_pcg, it_sor = st_pcg.stats['iterations'], st_sor.stats['iterations']
    print(case, 'field evaluations: pcg', it_pcg, 'sor', it_sor, flush=True)
    assert st_pcg.stats['converged'] and st_pcg.stats['epsilon'] < 1e-8
    assert 2 * it_pcg <= it_sor


def test_config1_single_pip_error(mb, oracle):
    """FP32 polynomials: the error against the FP64 oracle is reported per
    component; it stays at the 1e-3 level that FP32 Horner evaluation of
    these polynomials gives (their terms cancel by three orders of magnitude)."""
    waters, L = mb.systems.config1()
    sys = mb.forcefield.create_system(256, nonbonded_method='PME', box=L, components=1 | 2 | 4)
    pos = mb.forcefield.water_positions_with_m(waters)
    ctx = mb.Context(sys, precision='single-pip')
    ctx.setPositions(pos)
    st = ctx.getState(components=1 | 2 | 4)
    E, F = oracle.evaluate_system(sys, pos, components=1 | 2 | 4)
    got = st.getComponentEnergies()
    for k in ('twobody', 'threebody'):
        rel = abs(got[k] - E[k]) / abs(E[k])
        print('single-pip', k, '%.9g vs %.9g rel %.2e' % (got[k], E[k], rel), flush=True)
        assert rel <= 1e-2, k
    assert abs(got['onebody'] - E['onebody']) <= 1e-9 * abs(E['onebody'])
    err = rel_rms(st.getForces(), F)
    print('single-pip force rel rms err %.3e' % err, flush=True)
    assert err <= 1e-2


def test_lists_bit_exact_config1(mb, oracle):
    waters, L = mb.systems.config1()
    sys = mb.forcefield.create_system(256, nonbonded_method='PME', box=L, components=2 | 4)
    pos = mb.forcefield.water_positions_with_m(waters)
    ctx = mb.Context(sys, precision='double')
    ctx.setListMode(True)
    ctx.setPositions(pos)
    ctx.getState(components=2 | 4)
    opos = pos[0::4]
    want2 = oracle.dimer_list(opos, 0.65, True, [L] * 3)
    want3 = oracle.trimer_list(opos, 0.45, True, [L] * 3)
    got2, got3 = ctx.getDimerList(), ctx.getTrimerList()
    print('dimers', len(got2), len(want2), 'trimers', len(got3), len(want3), flush=True)
    assert sorted(map(tuple, got2.tolist())) == sorted(map(tuple, want2.tolist()))
    assert sorted(map(tuple, got3.tolist())) == sorted(map(tuple, want3.tolist()))


def test_lists_bit_exact_cluster(mb, oracle):
    waters = mb.systems.cluster(40, seed=7)
    sys = mb.forcefield.create_system(40, nonbonded_method='CutoffNonPeriodic', twobody_cutoff=0.65,
                                      threebody_cutoff=0.45, components=2 | 4)
    pos = mb.forcefield.water_positions_with_m(waters)
    ctx = mb.Context(sys, precision='double')
    ctx.setListMode(True)
    ctx.setPositions(pos)
    ctx.getState(components=2 | 4)
    opos = pos[0::4]
    assert sorted(map(tuple, ctx.getDimerList().tolist())) == sorted(
        map(tuple, oracle.dimer_list(opos, 0.65).tolist()))
    assert sorted(map(tuple, ctx.getTrimerList().tolist())) == sorted(
        map(tuple, oracle.trimer_list(opos, 0.45).tolist()))


@pytest.mark.parametrize('name,method,cutoff,box,comp', [('water3', 'CutoffNonPeriodic', 0.9, None, 31),
                                                         ('water50', 'PME', 0.9, 1.8, 31),
                                                         ('water_and_ion', 'CutoffNonPeriodic', 1.0, None, 8)])
def test_reference_fixture_systems(mb, oracle, name, method, cutoff, box, comp):
    pos, names, res, _ = mb.forcefield.read_pdb(os.path.join(GOLDEN, 'pdb', name + '.pdb'))
    nw = sum(1 for r in res if r == 'HOH') // 4
    sys = mb.forcefield.create_system(nw, nonbonded_method=method, nonbonded_cutoff=cutoff, box=box,
                                      ions=len(res) - 4 * nw, components=comp)
    compare(mb, oracle, sys, pos, 'double', components=comp, label=name)


def test_config2_properties(mb):
    """4096 waters at full size: deterministic, translation invariant, forces sum to ~0,
    mixed and FP64 paths agree."""
    waters, L = mb.systems.config2()
    sys = mb.forcefield.create_system(4096, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5,
                                      dipole_epsilon=1e-8, ewald_alpha=3.654826, pme_grid=(121, 121, 121))
    pos = mb.forcefield.water_positions_with_m(waters)
    out = {}
    for prec in ('double', 'mixed'):
        ctx = mb.Context(sys, precision=prec)
        ctx.setPositions(pos)
        a = ctx.getState()
        b = ctx.getState()
        # bitwise repeatable: every sum runs in an order fixed by the input (sorted rows,
        # bins and spreading table, fixed-point force accumulation)
        assert a.getPotentialEnergy() == b.getPotentialEnergy()
        assert np.array_equal(a.getForces(), b.getForces())
        ctx.setPositions(pos + np.array([0.37, -0.21, 1.3]))
        c = ctx.getState()
        # a rigid translation moves the atoms relative to the PME grid: the energy changes by
        # the reciprocal-space interpolation error only (Ewald tolerance 1e-5)
        assert abs(c.getPotentialEnergy() - a.getPotentialEnergy()) < 1e-5 * abs(a.getPotentialEnergy())
        f = a.getForces()
        assert np.all(np.isfinite(f))
        assert np.abs(f.sum(axis=0)).max() < 1e-3 * np.abs(f).max() * np.sqrt(len(f))
        out[prec] = a
        print(prec, a.getComponentEnergies(), a.stats, flush=True)
    ed, em = out['double'].getComponentEnergies(), out['mixed'].getComponentEnergies()
    tot = abs(out['double'].getPotentialEnergy())
    for k in ed:
        assert abs(ed[k] - em[k]) <= 1e-6 * tot, k
    assert rel_rms(out['mixed'].getForces(), out['double'].getForces()) <= 1e-4
    # the mixed mode solves by conjugate gradients, the FP64 mode by SOR
    assert out['mixed'].stats['iterations'] * 2 <= out['double'].stats['iterations']


def test_config2_ewald_split_invariance(mb):
    """4096 waters: the electrostatic energy does not depend on how it is split
    between real and reciprocal space.  rc 0.9 nm with the automatic alpha/grid
    for tolerance 1e-5 and rc 1.2 nm with its own automatic alpha/grid must agree
    to the accuracy that tolerance allows: the two splits differ by 3e-4 of the
    energy (the real-space truncation of the dipole terms dominates), checked at
    1e-3.  A missing neighbour or a wrong reciprocal term moves the energy by
    percents.  Both runs use the cell-list neighbour search."""
    waters, L = mb.systems.config2()
    pos = mb.forcefield.water_positions_with_m(waters)
    e = {}
    for rc in (0.9, 1.2):
        sys = mb.forcefield.create_system(4096, nonbonded_method='PME', box=L, nonbonded_cutoff=rc,
                                          ewald_error_tolerance=1e-5, dipole_epsilon=1e-8, components=16)
        ctx = mb.Context(sys, precision='double')
        ctx.setPositions(pos)
        st = ctx.getState(components=16)
        e[rc] = st.getComponentEnergies()['electrostatics']
        print('rc', rc, 'E_elec', e[rc], st.stats, flush=True)
    assert abs(e[0.9] - e[1.2]) <= 1e-3 * abs(e[0.9])


def test_pme_transform_paths_agree(mb, oracle):
    """The LDS grid transforms (every axis a product of radices 2..13) and the
    hipFFT transforms give the same PME electrostatics; an axis with a prime
    factor 17 selects hipFFT by itself, and that path matches the oracle too."""
    waters, L = mb.systems.config1()
    pos = mb.forcefield.water_positions_with_m(waters)
    sys = mb.forcefield.create_system(256, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5,
                                      dipole_epsilon=1e-8, components=16)
    res = {}
    for mode in ('lds', 'hipfft'):
        if mode == 'hipfft':
            os.environ['MBPOL_HIP_PME_FFT'] = 'hipfft'
        try:
            ctx = mb.Context(sys, precision='double')
        finally:
            os.environ.pop('MBPOL_HIP_PME_FFT', None)
        ctx.setPositions(pos)
        st = ctx.getState(components=16)
        res[mode] = (st.getComponentEnergies()['electrostatics'], st.getForces())
        print(mode, res[mode][0], st.stats, flush=True)
    assert abs(res['lds'][0] - res['hipfft'][0]) <= 1e-10 * abs(res['hipfft'][0])
    assert rel_rms(res['lds'][1], res['hipfft'][1]) <= 1e-10
    # 51 = 3 x 17: not an LDS length
    sys17 = mb.forcefield.create_system(256, nonbonded_method='PME', box=L, ewald_alpha=3.4, pme_grid=(48, 51, 50),
                                        dipole_epsilon=1e-8, components=16)
    compare(mb, oracle, sys17, pos, 'double', components=16, label='grid48x51x50')
