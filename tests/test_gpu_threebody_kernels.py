"""The 3-body kernels against each other and against the oracle on the
256-water box: the default kernel (the polynomial as 24 passes over the
trimer's relabelling group, csrc/pip_orbit_loop.h) and the term-by-term
kernel (MBPOL_HIP_3B_KERNEL=terms), both FP64.
They differ in summation order only (~1e-15 per trimer), but every trimer's
energy and forces are rounded to 2^-32 fixed point before they are summed, so
a rounding that flips moves the totals by 2^-32 kJ/mol per trimer: the bounds
between kernels are 1e-11 relative for the energy and 1e-10 for the RMS
force.  Against the oracle, the FP64 bar (1e-9 / 1e-8) applies.  The CPU
test tests/test_host_kernels.py checks the relabelling loop itself against
the reference's polynomial vectors."""
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


def threebody(mb, kernel, precision='double'):
    waters, L = mb.systems.config1()
    sysm = mb.forcefield.create_system(256, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5, components=4)
    old = os.environ.get('MBPOL_HIP_3B_KERNEL')
    os.environ['MBPOL_HIP_3B_KERNEL'] = kernel
    try:
        ctx = mb.Context(sysm, precision=precision)
    finally:
        if old is None:
            del os.environ['MBPOL_HIP_3B_KERNEL']
        else:
            os.environ['MBPOL_HIP_3B_KERNEL'] = old
    pos = mb.forcefield.water_positions_with_m(waters)
    ctx.setPositions(pos)
    st = ctx.getState(components=4)
    return sysm, pos, st.getPotentialEnergy(), st.getForces()


def test_threebody_kernels_agree_and_match_oracle(mb, oracle):
    sysm, pos, e0, f0 = threebody(mb, 'orbit')
    E, F = oracle.evaluate_system(sysm, pos, components=4)
    err = np.sqrt(np.mean((f0 - F) ** 2)) / np.sqrt(np.mean(F ** 2))
    print('orbit: 3-body %.12f vs oracle %.12f, force rel rms %.2e' % (e0, E['threebody'], err), flush=True)
    assert abs(e0 - E['threebody']) <= 1e-9 * abs(E['threebody'])
    assert err <= 1e-8
    for kernel in ('terms',):
        _, _, e1, f1 = threebody(mb, kernel)
        err = np.sqrt(np.mean((f1 - f0) ** 2)) / np.sqrt(np.mean(f0 ** 2))
        print('%s: 3-body %.12f vs orbit %.12f, force rel rms %.2e' % (kernel, e1, e0, err), flush=True)
        assert abs(e1 - e0) <= 1e-11 * abs(e0)
        assert err <= 1e-10
