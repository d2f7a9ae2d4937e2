"""The 4096-water periodic box (the benchmark workload) against the FP64
oracle's evaluation of the same coordinates, written by
tests/golden/make_config2_fixture.py into tests/golden/config2_oracle.npz
(the oracle takes about 8 minutes on one core, so it runs where the fixture
is made, not here).

* Both O-O lists, as sets, bit for bit: the 67,686 dimers within 0.65 nm and
  the 164,748 chain trimers (full reference chain list, i > j > k with
  d(i,j), d(j,k) <= 0.45 nm), compared through the count, the SHA-256 of the
  sorted rows and the rows per first molecule.
* The FP64 path (the reference's SOR solve) to 1e-9 of the total energy per
  component and 1e-8 relative RMS force; the mixed path (conjugate-gradient
  solve, FP32 dispersion pair terms) to 1e-6 and 1e-4.  Every component's
  error is printed.
"""
import hashlib
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), 'golden', 'config2_oracle.npz')
TOL = {'double': (1e-9, 1e-8), 'mixed': (1e-6, 1e-4)}


@pytest.fixture(scope='module')
def mb():
    import mbpol_openmm_plugin_amd as m
    if m.device_count() < 1:
        pytest.fail('no GPU visible to libmbpol_hip.so')
    return m


@pytest.fixture(scope='module')
def ref():
    return np.load(FIXTURE)


@pytest.fixture(scope='module')
def workload(mb, ref):
    waters, L = mb.systems.config2()
    sysm = mb.forcefield.create_system(4096, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5,
                                       dipole_epsilon=1e-8, ewald_alpha=3.654826, pme_grid=(121, 121, 121))
    pos = mb.forcefield.water_positions_with_m(waters)
    digest = hashlib.sha256(np.ascontiguousarray(pos, dtype=np.float64).tobytes()).hexdigest()
    assert digest == str(ref['positions_sha256']), 'the seeded generator gives other coordinates than the fixture'
    return sysm, pos


def digest(rows, nmol):
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    srt = np.ascontiguousarray(rows[np.lexsort(rows.T[::-1])])
    return len(rows), hashlib.sha256(srt.tobytes()).hexdigest(), np.bincount(rows[:, 0], minlength=nmol)


def test_config2_lists_bit_exact(mb, ref, workload):
    sysm, pos = workload
    ctx = mb.Context(sysm, precision='double')
    ctx.setListMode(True)  # the full chain list, triplets with a zero switch product included
    ctx.setPositions(pos)
    ctx.getState(components=2 | 4)
    for name, rows in (('dimer', ctx.getDimerList()), ('trimer', ctx.getTrimerList())):
        n, h, per = digest(rows, 4096)
        print(name, n, 'fixture', int(ref[name + '_count']), flush=True)
        bad = np.nonzero(per != ref[name + 's_per_first'])[0]
        assert len(bad) == 0, (name, 'row counts differ at first molecules', bad[:10])
        assert n == int(ref[name + '_count'])
        assert h == str(ref[name + '_sha256'])


@pytest.mark.parametrize('precision', ['double', 'mixed'])
def test_config2_against_oracle(mb, ref, workload, precision):
    sysm, pos = workload
    ctx = mb.Context(sysm, precision=precision)
    ctx.setPositions(pos)
    st = ctx.getState()
    got = st.getComponentEnergies()
    names = [str(c) for c in ref['components']]
    want = dict(zip(names, ref['energies']))
    etol, ftol = TOL[precision]
    scale = abs(want['total'])
    errs = {k: (got[k] - want[k]) / scale for k in names if k in got}
    errs['total'] = (st.getPotentialEnergy() - want['total']) / scale
    F = ref['forces']
    ferr = float(np.sqrt(np.mean((st.getForces() - F) ** 2)) / np.sqrt(np.mean(F ** 2)))
    print(precision, 'error / |E_total| per component:', {k: '%.2e' % v for k, v in errs.items()},
          'force rel rms %.2e' % ferr, 'iterations', st.stats['iterations'], 'oracle', int(ref['iterations']),
          flush=True)
    for k, v in errs.items():
        assert abs(v) <= etol, (precision, k, v)
    assert ferr <= ftol, (precision, ferr)
