"""Plane passes of the PME transforms (z and y passes fused per x plane, opt-in
with MBPOL_HIP_PME_PLANES=1) against the five single-axis passes, which the
parity tests pin to the FP64 oracle: the same electrostatic energy and forces
to 1e-11 relative (FP64 throughout; the two differ only in rounding), on the
256-water box and on the 4096-water box with its 121^3 grid.  The timer
counts show which passes ran.
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


def run(mb, waters, L, n, planes):
    sysm = mb.forcefield.create_system(n, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5,
                                       dipole_epsilon=1e-8, components=16)
    old = os.environ.get('MBPOL_HIP_PME_PLANES')
    os.environ['MBPOL_HIP_PME_PLANES'] = '1' if planes else '0'
    try:
        ctx = mb.Context(sysm, precision='double')
        ctx.setPositions(mb.forcefield.water_positions_with_m(waters))
        ctx.setTiming(True)
        st = ctx.getState(components=16)
        counts = {k: ctx.getTiming(k)[1] for k in ('fft_plane_fwd', 'fft_plane_inv', 'fft_z_r2c', 'fft_y')}
    finally:
        if old is None:
            del os.environ['MBPOL_HIP_PME_PLANES']
        else:
            os.environ['MBPOL_HIP_PME_PLANES'] = old
    return st.getPotentialEnergy(), st.getForces(), counts


@pytest.mark.parametrize('config', ['config1', 'config2'])
def test_plane_passes_match_single_axis_passes(mb, config):
    waters, L = getattr(mb.systems, config)()
    n = len(waters)
    e0, f0, c0 = run(mb, waters, L, n, False)
    e1, f1, c1 = run(mb, waters, L, n, True)
    print(config, 'energy %.10f vs %.10f' % (e1, e0), 'timers planes', c1, 'single-axis', c0, flush=True)
    assert c0['fft_plane_fwd'] == 0 and c0['fft_z_r2c'] > 0
    assert c1['fft_plane_fwd'] > 0 and c1['fft_plane_inv'] == c1['fft_plane_fwd'] and c1['fft_z_r2c'] == 0
    assert abs(e1 - e0) <= 1e-11 * abs(e0)
    err = np.sqrt(np.mean((f1 - f0) ** 2)) / np.sqrt(np.mean(f0 ** 2))
    print(config, 'force rel rms %.2e' % err, flush=True)
    assert err <= 1e-11
