"""The 2-body kernels against each other and against the oracle on the
256-water box: the polynomial as 8 passes over the dimer's relabelling group
-- swap the waters, swap the hydrogens and extra points of either water --
(csrc/pip_orbit_loop.h, MBPOL_HIP_2B_KERNEL=orbit) and the default
term-by-term kernel with its coefficients as instruction literals, both FP64.
They differ in summation order only (~1e-15 per dimer), but every dimer's
energy and forces are rounded to 2^-32 fixed point before they are summed, so
a rounding that flips moves the totals by 2^-32 kJ/mol per dimer: the bounds
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


def twobody(mb, kernel, precision='double'):
    waters, L = mb.systems.config1()
    sysm = mb.forcefield.create_system(256, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5, components=2)
    old = os.environ.get('MBPOL_HIP_2B_KERNEL')
    os.environ['MBPOL_HIP_2B_KERNEL'] = kernel
    try:
        ctx = mb.Context(sysm, precision=precision)
    finally:
        if old is None:
            del os.environ['MBPOL_HIP_2B_KERNEL']
        else:
            os.environ['MBPOL_HIP_2B_KERNEL'] = old
    assert ctx.getTuning('pip2_kernel') == (1 if kernel == 'terms' else 0)
    pos = mb.forcefield.water_positions_with_m(waters)
    ctx.setPositions(pos)
    st = ctx.getState(components=2)
    return sysm, pos, st.getPotentialEnergy(), st.getForces()


def test_twobody_kernels_agree_and_match_oracle(mb, oracle):
    sysm, pos, e0, f0 = twobody(mb, 'orbit')
    E, F = oracle.evaluate_system(sysm, pos, components=2)
    err = np.sqrt(np.mean((f0 - F) ** 2)) / np.sqrt(np.mean(F ** 2))
    print('orbit: 2-body %.12f vs oracle %.12f, force rel rms %.2e' % (e0, E['twobody'], err), flush=True)
    assert abs(e0 - E['twobody']) <= 1e-9 * abs(E['twobody'])
    assert err <= 1e-8
    _, _, e1, f1 = twobody(mb, 'terms')
    err = np.sqrt(np.mean((f1 - f0) ** 2)) / np.sqrt(np.mean(f0 ** 2))
    print('terms: 2-body %.12f vs orbit %.12f, force rel rms %.2e' % (e1, e0, err), flush=True)
    assert abs(e1 - e0) <= 1e-11 * abs(e0)
    assert err <= 1e-10


def test_twobody_single_precision_kernels(mb, oracle):
    """FP32 polynomials (single-pip precision): both kernels against the FP64
    oracle, at the 1e-3 level that FP32 evaluation of these polynomials gives
    (their terms cancel by three orders of magnitude; the two kernels sum in
    different orders, so they differ from each other at that level too)."""
    sysm, pos, e0, f0 = twobody(mb, 'orbit', 'single-pip')
    _, _, e1, f1 = twobody(mb, 'terms', 'single-pip')
    E, F = oracle.evaluate_system(sysm, pos, components=2)
    for name, e, f in (('orbit', e0, f0), ('terms', e1, f1)):
        rel = abs(e - E['twobody']) / abs(E['twobody'])
        err = np.sqrt(np.mean((f - F) ** 2)) / np.sqrt(np.mean(F ** 2))
        print('single-pip %s: %.8f vs oracle %.8f rel %.2e, force rel rms %.2e' % (name, e, E['twobody'], rel, err),
              flush=True)
        assert rel <= 1e-2 and err <= 1e-2
