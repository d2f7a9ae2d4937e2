"""The 3-body polynomial evaluated through its relabelling group (opt-in,
MBPOL_HIP_3B_SYM=1: E(x) = R(x) + sum over 24 relabellings of Q(y_h), see
csrc/gen_pip.py) against the term-by-term kernel, which the parity tests pin
to the FP64 oracle: the same 3-body energy and forces on the 256-water box.
Both are FP64 and differ in summation order only (~1e-15 per trimer), but
every trimer's energy and forces are rounded to 2^-32 fixed point before they
are summed, so a rounding that flips moves the totals by 2^-32 kJ/mol (per
trimer): the bounds are 1e-11 relative for the energy (5e-9 kJ/mol here)
and 1e-10 for the RMS force.  The CPU test
tests/test_pip_symmetry.py checks the decomposition itself on random
variables.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def mb():
    import mbpol_openmm_plugin_amd as m
    if m.device_count() < 1:
        pytest.fail('no GPU visible to libmbpol_hip.so')
    return m


def threebody(mb, sym):
    waters, L = mb.systems.config1()
    sysm = mb.forcefield.create_system(256, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5, components=4)
    old = os.environ.get('MBPOL_HIP_3B_SYM')
    os.environ['MBPOL_HIP_3B_SYM'] = '1' if sym else '0'
    try:
        ctx = mb.Context(sysm, precision='double')
    finally:
        if old is None:
            del os.environ['MBPOL_HIP_3B_SYM']
        else:
            os.environ['MBPOL_HIP_3B_SYM'] = old
    ctx.setPositions(mb.forcefield.water_positions_with_m(waters))
    st = ctx.getState(components=4)
    return st.getPotentialEnergy(), st.getForces()


def test_relabelling_group_matches_term_by_term(mb):
    e0, f0 = threebody(mb, False)
    e1, f1 = threebody(mb, True)
    err = np.sqrt(np.mean((f1 - f0) ** 2)) / np.sqrt(np.mean(f0 ** 2))
    print('3-body energy %.12f vs %.12f, force rel rms %.2e' % (e1, e0, err), flush=True)
    assert abs(e1 - e0) <= 1e-11 * abs(e0)
    assert err <= 1e-10
