"""Neighbour-row capacity of the electrostatics when it is checked after the
evaluation: a context that has sized its rows on the 256-water box is given a
configuration whose rows no longer fit (the waters moved rigidly towards the
box centre, so that the densest rows need more entries than the capacity).  That evaluation
overflows, is detected at its end and is run again with grown rows (the 2-
and 3-body work lists overflow too); its energies and forces must equal those
of a fresh context on the same configuration (same solver and start, 1e-12
relative).  For the FP64 path (SOR), the mixed path (conjugate gradients, whose
first attempt on truncated rows may stop without converging before the re-run)
and the mixed path in graph mode, where the overflow is found after a replay
and the evaluation is run again eagerly.
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


@pytest.mark.parametrize('mode', ['double', 'mixed', 'mixed-graph'])
def test_row_overflow_is_detected_and_redone(mb, mode):
    precision = mode.split('-')[0]
    waters, L = mb.systems.config1()
    sysm = mb.forcefield.create_system(256, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5,
                                       dipole_epsilon=1e-8)
    pos = mb.forcefield.water_positions_with_m(waters)
    tight = squeezed(pos, L, 0.82)
    a = mb.Context(sysm, precision=precision)
    a.setGraph(mode.endswith('graph'))
    a.setPositions(pos)
    for _ in range(3):
        a.getState()  # rows sized; later builds defer the capacity check (and in graph mode are captured)
    n0 = a.getStats()['num_elec_pairs']
    cap0 = a.getTuning('row_capacity_pol')
    a.setPositions(tight)
    sa = a.getState()
    cap1 = a.getTuning('row_capacity_pol')
    b = mb.Context(sysm, precision=precision)
    b.setPositions(tight)
    sb = b.getState()
    n1, n2 = a.getStats()['num_elec_pairs'], b.getStats()['num_elec_pairs']
    print('site pairs: spread box %d, squeezed %d (fresh context %d); row capacity %d -> %d; energies %.10f %.10f' %
          (n0, n1, n2, cap0, cap1, sa.getPotentialEnergy(), sb.getPotentialEnergy()), flush=True)
    assert n1 == n2 and n1 > n0
    assert cap1 > cap0  # the squeezed rows overflowed and grew
    if mode.endswith('graph'):
        print('graph: captures %d, replays %d, fallbacks %d' % tuple(
            a.getTuning(k) for k in ('graph_captures', 'graph_replays', 'graph_fallbacks')), flush=True)
        assert a.getTuning('graph_fallbacks') >= 1
    assert abs(sa.getPotentialEnergy() - sb.getPotentialEnergy()) <= 1e-12 * abs(sb.getPotentialEnergy())
    fa, fb = sa.getForces(), sb.getForces()
    assert np.sqrt(np.mean((fa - fb) ** 2)) <= 1e-12 * np.sqrt(np.mean(fb ** 2))
