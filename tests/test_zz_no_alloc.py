"""mbpol_context_get_electrostatic_potential_noalloc (no counterpart in the
reference; for callers that cannot allocate) against the allocating entry
point mbpol_context_get_electrostatic_potential, through the C ABI: the same
kernel on the same charges and dipoles, so the results must be identical, on
two calls in a row, the second with fewer points (the fixed scratch array keeps
the first call's contents).  A request above the array's capacity (65536
points) is refused before any GPU work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def mb():
    import mbpol_openmm_plugin_amd as m
    if m.device_count() < 1:
        pytest.fail('no GPU visible to libmbpol_hip.so')
    return m


def call(mb, fn, ctx, pts):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    out = np.full(len(pts), np.nan)
    rc = fn(ctx._h, len(pts), pts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return rc, out


@pytest.mark.parametrize('which', ['cluster', 'box'])
def test_noalloc_potential_matches_allocating_entry_point(mb, which):
    if which == 'cluster':
        waters = mb.systems.config0()
        sysm = mb.forcefield.create_system(14, nonbonded_method='CutoffNonPeriodic', twobody_cutoff=1e10,
                                           threebody_cutoff=1e10, nonbonded_cutoff=1e3, dipole_epsilon=1e-8)
        box = 2.0
    else:
        waters, box = mb.systems.config1()
        sysm = mb.forcefield.create_system(256, nonbonded_method='PME', box=box, ewald_error_tolerance=1e-5,
                                           dipole_epsilon=1e-8)
    pos = mb.forcefield.water_positions_with_m(waters)
    ctx = mb.Context(sysm, precision='double')
    ctx.setPositions(pos)
    ctx.getState(components=mb.api.MBPOL_ELECTROSTATICS)
    L = mb.api.library()
    rng = np.random.default_rng(3)
    for npts in (5000, 1234):  # two calls in a row, the second smaller
        pts = rng.uniform(-0.5, box + 0.5, (npts, 3))
        rc0, want = call(mb, L.mbpol_context_get_electrostatic_potential, ctx, pts)
        rc1, got = call(mb, L.mbpol_context_get_electrostatic_potential_noalloc, ctx, pts)
        assert rc0 == 0 and rc1 == 0, mb.api.library().mbpol_last_error()
        print(which, npts, 'points, max |diff|', np.abs(got - want).max(), flush=True)
        assert np.array_equal(got, want)
    rc, _ = call(mb, L.mbpol_context_get_electrostatic_potential_noalloc, ctx, np.zeros((65537, 3)))
    assert rc == 1
