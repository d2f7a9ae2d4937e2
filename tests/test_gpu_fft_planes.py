"""Plane passes of the PME transforms (z and y passes fused per x plane, opt-in
with MBPOL_HIP_PME_PLANES=1) against the five single-axis passes, which the
parity tests pin to the FP64 oracle: the same electrostatic energy and forces
to 1e-11 relative (FP64 throughout; the two differ only in rounding), on the
256-water box and on the 4096-water box with its 121^3 grid.  The timer
counts show which passes ran.

The dipole spreading fused with the z pass (the default) against the separate
spreading kernel and z pass (MBPOL_HIP_SPREAD_FUSED=0): the fused kernel sums
the same rows of the spreading table in the same order, so energies, forces,
dipoles and iteration counts agree to 1e-13 relative (equal unless the
compiler contracts the row sums differently in the two kernels).
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


def run(mb, waters, L, n, planes, fused=True, dipoles=False):
    sysm = mb.forcefield.create_system(n, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5,
                                       dipole_epsilon=1e-8, components=16)
    env = {'MBPOL_HIP_PME_PLANES': '1' if planes else '0', 'MBPOL_HIP_SPREAD_FUSED': '1' if fused else '0'}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        ctx = mb.Context(sysm, precision='double')
        ctx.setPositions(mb.forcefield.water_positions_with_m(waters))
        ctx.setTiming(True)
        st = ctx.getState(components=16)
        counts = {k: ctx.getTiming(k)[1] for k in ('fft_plane_fwd', 'fft_plane_inv', 'fft_z_r2c', 'fft_y',
                                                   'pme_spread', 'spread_fft_z')}
        extra = (ctx.getInducedDipoles(), ctx.getStats()['iterations']) if dipoles else None
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    return st.getPotentialEnergy(), st.getForces(), counts, extra


@pytest.mark.parametrize('config', ['config1', 'config2'])
def test_plane_passes_match_single_axis_passes(mb, config):
    waters, L = getattr(mb.systems, config)()
    n = len(waters)
    e0, f0, c0, _ = run(mb, waters, L, n, False)
    e1, f1, c1, _ = run(mb, waters, L, n, True)
    print(config, 'energy %.10f vs %.10f' % (e1, e0), 'timers planes', c1, 'single-axis', c0, flush=True)
    assert c0['fft_plane_fwd'] == 0 and c0['fft_z_r2c'] > 0
    assert c1['fft_plane_fwd'] > 0 and c1['fft_plane_inv'] == c1['fft_plane_fwd'] and c1['fft_z_r2c'] == 0
    assert abs(e1 - e0) <= 1e-11 * abs(e0)
    err = np.sqrt(np.mean((f1 - f0) ** 2)) / np.sqrt(np.mean(f0 ** 2))
    print(config, 'force rel rms %.2e' % err, flush=True)
    assert err <= 1e-11


@pytest.mark.parametrize('config', ['config1', 'config2'])
def test_fused_spreading_matches_separate_passes(mb, config):
    waters, L = getattr(mb.systems, config)()
    n = len(waters)
    e0, f0, c0, (d0, i0) = run(mb, waters, L, n, False, fused=False, dipoles=True)
    e1, f1, c1, (d1, i1) = run(mb, waters, L, n, False, fused=True, dipoles=True)
    print(config, 'energy %.12f vs %.12f' % (e1, e0), 'iterations', i1, i0, 'timers fused', c1, 'separate', c0,
          flush=True)
    assert c0['spread_fft_z'] == 0 and c0['pme_spread'] > 0
    # the charge grid keeps its own spreading and z pass; every dipole grid goes through the fused kernel
    assert c1['spread_fft_z'] > 0 and c1['pme_spread'] == 0
    assert c1['spread_fft_z'] + c1['fft_z_r2c'] == c0['fft_z_r2c'] and c1['spread_fft_z'] == c0['pme_spread']
    assert i1 == i0
    assert abs(e1 - e0) <= 1e-13 * abs(e0)
    ferr = np.sqrt(np.mean((f1 - f0) ** 2)) / np.sqrt(np.mean(f0 ** 2))
    derr = np.sqrt(np.mean((d1 - d0) ** 2)) / np.sqrt(np.mean(d0 ** 2))
    print(config, 'force rel rms %.2e, dipole rel rms %.2e, bit-equal %s' %
          (ferr, derr, np.array_equal(f1, f0) and np.array_equal(d1, d0)), flush=True)
    assert ferr <= 1e-13 and derr <= 1e-13
