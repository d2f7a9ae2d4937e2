"""Neighbour-row capacity of the electrostatics when it is checked after the
evaluation: a context that has sized its rows on the 256-water box is given a
configuration whose rows no longer fit (the waters moved rigidly towards the
box centre, so that the densest rows need more entries than the capacity).  That evaluation
overflows, is detected at its end and is run again with grown rows; its
energies and forces must equal those of a fresh context on the same
configuration (FP64 path, same solver and start, 1e-12 relative).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def mb():
    import mbpol_openmm_plugin_amd as m
    if m.device_count() < 1:
        pytest.fail('no GPU visible to libmbpol_hip.so')
    return m


def squeezed(pos, L, f):
    """Waters moved rigidly: oxygen-to-centre vectors scaled by f (sites of a water keep their offsets)."""
    out = pos.copy()
    c = np.full(3, L / 2)
    for w in range(len(pos) // 4):
        o = pos[4 * w]
        shift = (c + (o - c) * f) - o
        out[4 * w:4 * w + 4] += shift
    return out


def test_row_overflow_is_detected_and_redone(mb):
    waters, L = mb.systems.config1()
    sysm = mb.forcefield.create_system(256, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5,
                                       dipole_epsilon=1e-8)
    pos = mb.forcefield.water_positions_with_m(waters)
    tight = squeezed(pos, L, 0.82)
    a = mb.Context(sysm, precision='double')
    a.setPositions(pos)
    a.getState()
    a.getState()  # rows sized; later builds defer the capacity check
    n0 = a.getStats()['num_elec_pairs']
    cap0 = a.getTuning('row_capacity_pol')
    a.setPositions(tight)
    sa = a.getState()
    cap1 = a.getTuning('row_capacity_pol')
    b = mb.Context(sysm, precision='double')
    b.setPositions(tight)
    sb = b.getState()
    n1, n2 = a.getStats()['num_elec_pairs'], b.getStats()['num_elec_pairs']
    print('site pairs: spread box %d, squeezed %d (fresh context %d); row capacity %d -> %d; energies %.10f %.10f' %
          (n0, n1, n2, cap0, cap1, sa.getPotentialEnergy(), sb.getPotentialEnergy()), flush=True)
    assert n1 == n2 and n1 > n0
    assert cap1 > cap0  # the squeezed rows overflowed and grew
    assert abs(sa.getPotentialEnergy() - sb.getPotentialEnergy()) <= 1e-12 * abs(sb.getPotentialEnergy())
    fa, fb = sa.getForces(), sb.getForces()
    assert np.sqrt(np.mean((fa - fb) ** 2)) <= 1e-12 * np.sqrt(np.mean(fb ** 2))
